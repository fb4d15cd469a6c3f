// Sparse VFE GPR (Sgpr) on one GPU, in gprx's whitened formulation
// (src/sgpr/factor/vfe.rs:191-369, src/sgpr/factor/derivatives.rs:94-201):
//   L = chol(K_mm + j I), X = L^-1, U = K_mn           (m x n)
//   A = X U                                           (TRMM on MFMA)
//   B = A A^T + s2 I                                  (split-K SYRK on MFMA)
//   w = B^-1 A y, r = y - A^T w
//   F = 1/2 [n log 2pi + (n - m) log s2 + log|B| + (|y|^2 - (A y)'w)/s2]
//       + (sum k_ii - |A|_F^2) / (2 s2)
// The adjoint G = B^-1 A - (w r^T + A)/s2 is never formed: with
// Q = X^T (B^-1 - I/s2) and v = X^T w,
//   w_mn = X^T G = Q A - v r^T / s2      (Q A on MFMA, rank-1 part in the contraction)
//   G A^T = I - s2 B^-1 - (B - s2 I)/s2 - w (A r)^T / s2 (+ I)   (m x m only)
//   w_mm = -sym(X^T tril_half(G A^T) X)
//   dF/dtheta = <w_mm, dK_mm> + <w_mn, dK_mn> + 1/(2 s2) sum dk_ii
//   dF/dlog s2 = 1/2 ((n - m) + s2 tr B^-1 + |w|^2 - quad) - trace
// so the data-sized products are 4 m^2 n (A, A A^T, Q A) against gprx's
// 7 m^2 n.  The m x m algebra and both Cholesky factors are f64; U, A, A A^T
// and Q A run in f32 on MFMA in the mixed form.
#include <cstring>
#include <memory>
#include <type_traits>

#include "model.hpp"

namespace gprx {

int exact_validate_theta(const HostProgram& hp, const double* th, size_t p, gprx_error* err);
void exact_upload_points(DevBuf<double>& dst, const double* src, int n, int d, long rows_pad, int dp,
                         hipStream_t s);

}  // namespace gprx

using namespace gprx;

namespace {

constexpr double kLog2Pi = 1.8378770664093453;

template <typename F>
int guarded(gprx_error* err, F&& f) {
  try {
    return f();
  } catch (const DeviceError& e) {
    set_error(err, GPRX_DEVICE_ERROR, e.msg);
    return GPRX_DEVICE_ERROR;
  } catch (...) {
    set_error(err, GPRX_DEVICE_ERROR, "unexpected failure");
    return GPRX_DEVICE_ERROR;
  }
}

struct SgprModel : gprx_hip_model {
  SgprModel() : gprx_hip_model(kSgpr) {}
  ~SgprModel() override {
    if (stream) (void)hipStreamSynchronize(stream);
    fstreams.destroy();
    if (stream) {
      for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
      (void)hipStreamDestroy(stream);
    }
  }
  HostProgram hp;
  int n = 0, m = 0, d = 0, dp = 0, mt = 0, nchunks = 1;
  long npad = 0, mpad = 0, ldt = 0, kc = 0;
  int precision = GPRX_FP64;
  std::vector<double> theta;
  double ynorm2 = 0.0, kmm_jitter = 0.0;
  bool factored = false;
  hipStream_t stream = nullptr;
  FactorStreams fstreams;
  hipEvent_t ev[2] = {nullptr, nullptr};
  DevBuf<double> x, y, z;
  DevBuf<DevProgram> prog;
  HostBuf<DevProgram> hprog;
  // m x m, f64
  DevBuf<double> KA, Xmm, Tm, S, BA, XB, Dm, Q64, H, T1, T2, Wmm;
  // m x n and split-K partials, storage type
  DevBuf<double> U64, A64, Sp64, Xc64;
  DevBuf<float> U32, A32, Sp32, Xc32, Q32;
  DevBuf<double> logd_mm, logd_B, b, w, v, tv, r, ur, Ar, gpart, partials, partials2, accs, scal;
  DevBuf<int> info;
  HostBuf<double> hscal, hacc;
  HostBuf<int> hinfo;

  DevPoints zp() const { return DevPoints{z.p, mpad, m}; }
  DevPoints xp() const { return DevPoints{x.p, npad, n}; }
  bool mixed() const { return precision == GPRX_MIXED; }

  void upload_program(const double* th, double jitter) {
    make_dev_program(hp, th, jitter, hprog.p);
    GPRX_HIP_CHECK(hipMemcpyAsync(prog.p, hprog.p, sizeof(DevProgram), hipMemcpyHostToDevice, stream));
  }

  int sync_info() {
    GPRX_HIP_CHECK(hipMemcpyAsync(hinfo.p, info.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    GPRX_HIP_CHECK(hipStreamSynchronize(stream));
    return *hinfo.p;
  }

  // K_mm + j I, its factor and inverse.  Returns info bits.
  int factor_kmm(const double* th, double j) {
    hipStream_t s = stream;
    upload_program(th, j);
    GPRX_HIP_CHECK(hipMemsetAsync(info.p, 0, sizeof(int), s));
    fill_lower<double>(dp, prog.p, zp(), KA.p, mpad, nullptr, info.p, s);
    factor_inverse<double>(KA.p, Xmm.p, Tm.p, mpad, ldt, mt, logd_mm.p, info.p, fstreams);
    return sync_info();
  }

  // U = K(Z, X), A = X U, S = A A^T (lower), b = A y.
  template <typename T>
  void data_pass(T* U, T* A, T* Xc, T* Sp) {
    hipStream_t s = stream;
    const int M = (int)mpad, N = (int)npad;
    fill_cross<T>(dp, prog.p, zp(), xp(), U, mpad, s);
    if (std::is_same<T, double>::value)
      GPRX_HIP_CHECK(hipMemcpyAsync(Xc, Xmm.p, sizeof(double) * mpad * mpad, hipMemcpyDeviceToDevice, s));
    else
      convert<float, double>(reinterpret_cast<float*>(Xc), Xmm.p, (size_t)mpad * mpad, s);
    gemm_call<T>(false, false, Xc, mpad, U, mpad, A, mpad, M, N, M, 1.0, 0.0, kKhiRow, 0,
                 kOrderColFastRev, s);
    GemmArgs<T> g{};
    g.A = A;
    g.lda = mpad;
    g.B = A;
    g.ldb = mpad;
    g.C = Sp;
    g.ldc = mpad;
    g.M = g.N = M;
    g.K = (int)kc;
    g.alpha = 1.0;
    g.beta = 0.0;
    g.lower = 1;
    g.batch = nchunks;
    g.strideA = g.strideB = kc * mpad;
    g.strideC = mpad * mpad;
    gemm<T>(false, true, g, s);
    sum_partials_lower<T>(Sp, nchunks, mpad, M, S.p, s);
    gemv_n<T>(A, mpad, M, N, y.p, b.p, gpart.p, 0, 0, 0, s);
  }

  // r = y - A^T w, Ar = A r.
  template <typename T>
  void resid_pass(const T* A) {
    gemv_t<T>(A, mpad, (int)mpad, (int)npad, w.p, ur.p, 0, 0, 0, 0, stream);
    vec_combine(r.p, n, (int)npad, 1.0, y.p, -1.0, ur.p, stream);
    gemv_n<T>(A, mpad, (int)mpad, (int)npad, r.p, Ar.p, gpart.p, 0, 0, 0, stream);
  }

  int evaluate(const double* th, bool want_grad, double* value, double* grad, gprx_error* err) {
    hipStream_t s = stream;
    const int P = hp.n_params;
    const double s2 = std::exp(th[P]);
    const int M = (int)mpad;
    // K_mm with the sparse default adaptive jitter (src/policy.rs:222-226).
    double j = 0.0, next = 1e-8;
    int remaining = 5;
    for (;;) {
      const int bits = factor_kmm(th, j);
      if (bits == 0) break;
      factored = false;
      if (bits & 2) {
        set_error(err, GPRX_NON_FINITE_KERNEL_VALUE, "K_mm holds a non-finite value");
        return GPRX_NON_FINITE_KERNEL_VALUE;
      }
      if (remaining == 0 || next > 1e-3) {
        set_error(err, GPRX_CHOLESKY_FAILED, "K_mm Cholesky failed", GPRX_STAGE_FIT, j, (size_t)m);
        return GPRX_CHOLESKY_FAILED;
      }
      --remaining;
      j = next;
      next *= 10.0;
    }
    kmm_jitter = j;
    GPRX_HIP_CHECK(hipEventRecord(ev[0], s));
    if (mixed())
      data_pass<float>(U32.p, A32.p, Xc32.p, Sp32.p);
    else
      data_pass<double>(U64.p, A64.p, Xc64.p, Sp64.p);
    // B = A A^T + s2 I; no jitter retries (src/sgpr/factor/vfe.rs:244-250).
    symmetrize_lower(S.p, mpad, M, s);
    mat_combine(BA.p, mpad, M, M, 1.0, S.p, mpad, 0.0, nullptr, 0, s);
    add_diag(BA.p, mpad, M, s2, s);
    GPRX_HIP_CHECK(hipMemsetAsync(info.p, 0, sizeof(int), s));
    factor_inverse<double>(BA.p, XB.p, Tm.p, mpad, ldt, mt, logd_B.p, info.p, fstreams);
    gemv_n<double>(XB.p, mpad, M, M, b.p, tv.p, gpart.p, 1, 0, 0, s);
    gemv_t<double>(XB.p, mpad, M, M, tv.p, w.p, 1, 0, 0, 0, s);
    // B^-1 into BA (consumed by the factor), full storage.
    gemm_call<double>(true, false, XB.p, mpad, XB.p, mpad, BA.p, mpad, M, M, M, 1.0, 0.0, kKloRow,
                      1, kOrderRowFast, s);
    symmetrize_lower(BA.p, mpad, M, s);
    if (mixed())
      resid_pass<float>(A32.p);
    else
      resid_pass<double>(A64.p);
    // scalars: [0] (Ay)'w, [1] sum log L_B, [2] |w|^2, [3] -, [4] |A|_F^2,
    // [5] tr B^-1, [7] k(x, x)
    dot_and_sum(b.p, w.p, logd_B.p, m, scal.p, s);
    dot_and_sum(w.p, w.p, w.p, m, scal.p + 2, s);
    diag_sum(S.p, mpad, m, scal.p + 4, s);
    diag_sum(BA.p, mpad, m, scal.p + 5, s);
    fill_diag(dp, prog.p, DevPoints{z.p, mpad, 1}, scal.p + 7, s);
    if (want_grad) {
      // Q = X^T (B^-1 - I/s2), v = X^T w
      mat_combine(Dm.p, mpad, M, M, 1.0, BA.p, mpad, 0.0, nullptr, 0, s);
      add_diag(Dm.p, mpad, M, -1.0 / s2, s);
      gemm_call<double>(true, false, Xmm.p, mpad, Dm.p, mpad, Q64.p, mpad, M, M, M, 1.0, 0.0,
                        kKloRow, 0, kOrderRowFast, s);
      gemv_t<double>(Xmm.p, mpad, M, M, w.p, v.p, 1, 0, 0, 0, s);
      // w_mn = Q A - v r^T / s2: Q A on MFMA into U's storage (K_mn is not
      // needed once A is formed), the rank-1 part inside the contraction.
      const int tiles1 = (int)(mpad / kTile) * (int)(npad / kTile);
      const int N = (int)npad;
      if (mixed()) {
        convert<float, double>(Q32.p, Q64.p, (size_t)mpad * mpad, s);
        gemm_call<float>(false, false, Q32.p, mpad, A32.p, mpad, U32.p, mpad, M, N, M, 1.0, 0.0, 0,
                         0, kOrderRowFast, s);
        contract_cross<float>(dp, hp.ns, prog.p, zp(), xp(), U32.p, mpad, partials.p, s, v.p, r.p, 1.0 / s2);
      } else {
        gemm_call<double>(false, false, Q64.p, mpad, A64.p, mpad, U64.p, mpad, M, N, M, 1.0, 0.0, 0,
                          0, kOrderRowFast, s);
        contract_cross<double>(dp, hp.ns, prog.p, zp(), xp(), U64.p, mpad, partials.p, s, v.p, r.p,
                               1.0 / s2);
      }
      reduce_partials(partials.p, tiles1, accs.p, s);
      // G A^T = I - s2 B^-1 - S/s2 - w (A r)^T / s2 ; w_mm = -sym(X^T tril_half(.) X)
      mat_combine(H.p, mpad, M, M, -s2, BA.p, mpad, -1.0 / s2, S.p, mpad, s);
      add_diag(H.p, mpad, M, 1.0, s);
      outer_add(H.p, mpad, M, M, -1.0 / s2, w.p, Ar.p, s);
      tril_half<double>(H.p, mpad, M, T2.p, s);
      gemm_call<double>(true, false, Xmm.p, mpad, T2.p, mpad, T1.p, mpad, M, M, M, 1.0, 0.0,
                        kKloRow, 0, kOrderRowFast, s);
      gemm_call<double>(false, false, T1.p, mpad, Xmm.p, mpad, T2.p, mpad, M, M, M, 1.0, 0.0,
                        kKloCol, 0, kOrderRowFast, s);
      neg_sym(T2.p, mpad, M, Wmm.p, s);
      contract_cross<double>(dp, hp.ns, prog.p, zp(), zp(), Wmm.p, mpad, partials2.p, s);
      reduce_partials(partials2.p, cross_tiles(mpad, mpad), accs.p + kMaxAcc, s);
      contract_diag(dp, prog.p, xp(), 0.5 / s2, accs.p + 2 * kMaxAcc, s);
    }
    GPRX_HIP_CHECK(hipEventRecord(ev[1], s));
    GPRX_HIP_CHECK(hipMemcpyAsync(hscal.p, scal.p, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
    if (want_grad)
      GPRX_HIP_CHECK(hipMemcpyAsync(hacc.p, accs.p, sizeof(double) * 3 * kMaxAcc, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(hinfo.p, info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
    last_eval_ms = ms;
    if (*hinfo.p) {
      factored = false;
      set_error(err, GPRX_CHOLESKY_FAILED, "Cholesky of A A^T + s2 I failed", GPRX_STAGE_FIT, 0.0,
                (size_t)m);
      return GPRX_CHOLESKY_FAILED;
    }
    const double* sc = hscal.p;
    const double ayw = sc[0], logdet_B = 2.0 * sc[1], w2 = sc[2], af2 = sc[4], trBinv = sc[5];
    const double kdiag_sum = (double)n * sc[7];
    const double quad = (ynorm2 - ayw) / s2;
    const double trace = (kdiag_sum - af2) / (2.0 * s2);
    const double nm = (double)n - (double)m;
    *value = 0.5 * ((double)n * kLog2Pi + nm * std::log(s2) + logdet_B + quad) + trace;
    if (!std::isfinite(*value)) {
      factored = false;
      set_error(err, GPRX_NON_FINITE_KERNEL_VALUE, "non-finite bound");
      return GPRX_NON_FINITE_KERNEL_VALUE;
    }
    if (want_grad) {
      double acc[kMaxAcc];
      for (int k = 0; k < kMaxAcc; ++k) acc[k] = hacc.p[k] + hacc.p[kMaxAcc + k] + hacc.p[2 * kMaxAcc + k];
      acc_to_grad(hp, th, acc, grad);
      grad[P] = 0.5 * (nm + s2 * trBinv + w2 - quad) - trace;
    }
    factored = true;
    return GPRX_OK;
  }

  // src/sgpr/factor/predict.rs:102-260: a = X k_m*, mean = a'w,
  // latent = k** - |a|^2 + s2 |L_B^-1 a|^2.
  int predict(const double* xs, size_t q, int32_t kind, double* mean, double* var, gprx_error* err) {
    if (!factored) {
      double vv;
      int rc = evaluate(theta.data(), false, &vv, nullptr, err);
      if (rc != GPRX_OK) {
        if (err) err->stage = GPRX_STAGE_PREDICT;
        return rc;
      }
    } else {
      upload_program(theta.data(), kmm_jitter);
    }
    hipStream_t s = stream;
    const long qpad = (long)round_up(q, kTile);
    DevBuf<double> xq, Ks, V1, V2, mu, s1, s2v, kss;
    exact_upload_points(xq, xs, (int)q, d, qpad, dp, s);
    DevPoints pq{xq.p, qpad, (int)q};
    Ks.alloc((size_t)mpad * qpad);
    V1.alloc((size_t)mpad * qpad);
    V2.alloc((size_t)mpad * qpad);
    mu.alloc(qpad);
    s1.alloc(qpad);
    s2v.alloc(qpad);
    kss.alloc(qpad);
    fill_cross<double>(dp, prog.p, zp(), pq, Ks.p, mpad, s);
    gemm_call<double>(false, false, Xmm.p, mpad, Ks.p, mpad, V1.p, mpad, (int)mpad, (int)qpad,
                      (int)mpad, 1.0, 0.0, kKhiRow, 0, kOrderColFastRev, s);
    gemv_t<double>(V1.p, mpad, (int)mpad, (int)qpad, w.p, mu.p, 0, 0, 0, 0, s);
    gemm_call<double>(false, false, XB.p, mpad, V1.p, mpad, V2.p, mpad, (int)mpad, (int)qpad,
                      (int)mpad, 1.0, 0.0, kKhiRow, 0, kOrderColFastRev, s);
    col_sumsq<double>(V1.p, mpad, (int)mpad, (int)qpad, s1.p, s);
    col_sumsq<double>(V2.p, mpad, (int)mpad, (int)qpad, s2v.p, s);
    fill_diag(dp, prog.p, pq, kss.p, s);
    std::vector<double> h1(q), h2(q), hk(q);
    GPRX_HIP_CHECK(hipMemcpyAsync(mean, mu.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(h1.data(), s1.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(h2.data(), s2v.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(hk.data(), kss.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    if (var) {
      const double noise = std::exp(theta[hp.n_params]);
      for (size_t k = 0; k < q; ++k) {
        double lat = hk[k] - h1[k] + noise * h2[k];
        if (lat < 0.0) lat = 0.0;
        var[k] = kind == GPRX_VARIANCE_LATENT ? lat : lat + noise;
      }
    }
    return GPRX_OK;
  }
};

}  // namespace

extern "C" {

int gprx_hip_sgpr_create(const double* x, size_t n, size_t d, const double* y, const double* z,
                         size_t m, const gprx_kernel_program* kernel, const double* log_theta,
                         size_t p, int32_t precision, gprx_hip_model** out, gprx_error* err) {
  if (out) *out = nullptr;
  return guarded(err, [&]() -> int {
    if (!x || !y || !z || !out || n == 0 || d == 0 || m == 0) {
      set_error(err, GPRX_EMPTY_INPUT, "empty data or inducing set");
      return GPRX_EMPTY_INPUT;
    }
    if (d > (size_t)kMaxDim) {
      set_error(err, GPRX_UNSUPPORTED, "input dimension above 32");
      return GPRX_UNSUPPORTED;
    }
    if (m > 16384 || n > ((size_t)1 << 26)) {
      set_error(err, GPRX_UNSUPPORTED, "sparse sizes above m = 16384 or n = 2^26");
      return GPRX_UNSUPPORTED;
    }
    if (!all_finite(x, n * d) || !all_finite(y, n) || !all_finite(z, m * d)) {
      set_error(err, GPRX_NON_FINITE_INPUT, "non-finite data");
      return GPRX_NON_FINITE_INPUT;
    }
    if (precision != GPRX_FP64 && precision != GPRX_MIXED) {
      set_error(err, GPRX_UNSUPPORTED, "unknown precision");
      return GPRX_UNSUPPORTED;
    }
    auto mo = std::make_unique<SgprModel>();
    std::string why;
    int rc = parse_program(kernel, (int)d, &mo->hp, &why);
    if (rc != GPRX_OK) {
      set_error(err, rc, why);
      return rc;
    }
    if (!log_theta) {
      set_error(err, GPRX_LENGTH_MISMATCH, "missing parameters");
      return GPRX_LENGTH_MISMATCH;
    }
    rc = exact_validate_theta(mo->hp, log_theta, p, err);
    if (rc != GPRX_OK) return rc;
    SgprModel& M = *mo;
    M.n = (int)n;
    M.m = (int)m;
    M.d = (int)d;
    M.dp = padded_dim((int)d);
    M.mpad = (long)round_up(m, kTile);
    M.mt = (int)(M.mpad / kTile);
    M.ldt = (long)((M.mt + 1) / 2) * kTile;
    // Split the n-long reduction of U U^T into chunks so the launch holds
    // at most 2048 workgroups (four per slot at two per CU: a launch just
    // above a multiple of 512 ends on a wave of a few); pad n to a whole
    // number of chunks.
    const long ntiles = (long)round_up(n, kTile) / kTile;
    const long lower = (long)M.mt * (M.mt + 1) / 2;
    long want = std::max<long>(1, std::min<long>(ntiles, 2048 / lower));
    M.kc = ((ntiles + want - 1) / want) * kTile;
    M.npad = (long)round_up(n, (size_t)M.kc);
    M.nchunks = (int)(M.npad / M.kc);
    M.precision = precision;
    M.theta.assign(log_theta, log_theta + p);
    M.ynorm2 = 0.0;
    for (size_t i = 0; i < n; ++i) M.ynorm2 += y[i] * y[i];
    GPRX_HIP_CHECK(hipStreamCreateWithFlags(&M.stream, hipStreamNonBlocking));
    M.fstreams.init(M.stream);
    for (auto& e : M.ev) GPRX_HIP_CHECK(hipEventCreate(&e));
    hipStream_t s = M.stream;
    exact_upload_points(M.x, x, (int)n, (int)d, M.npad, M.dp, s);
    exact_upload_points(M.z, z, (int)m, (int)d, M.mpad, M.dp, s);
    {
      std::vector<double> yh(M.npad, 0.0);
      std::memcpy(yh.data(), y, sizeof(double) * n);
      M.y.alloc(M.npad);
      GPRX_HIP_CHECK(hipMemcpyAsync(M.y.p, yh.data(), sizeof(double) * M.npad, hipMemcpyHostToDevice, s));
    }
    const size_t mm = (size_t)M.mpad * M.mpad;
    for (DevBuf<double>* b : {&M.KA, &M.Xmm, &M.S, &M.BA, &M.XB, &M.Dm, &M.Q64, &M.H, &M.T1, &M.T2, &M.Wmm}) {
      b->alloc(mm);
      GPRX_HIP_CHECK(hipMemsetAsync(b->p, 0, sizeof(double) * mm, s));
    }
    M.Tm.alloc((size_t)M.ldt * factor_tmp_cols(M.mt) * kTile);
    const size_t un = (size_t)M.mpad * M.npad;
    if (precision == GPRX_MIXED) {
      M.U32.alloc(un);
      M.A32.alloc(un);
      M.Sp32.alloc(mm * M.nchunks);
      M.Xc32.alloc(mm);
      M.Q32.alloc(mm);
    } else {
      M.U64.alloc(un);
      M.A64.alloc(un);
      M.Sp64.alloc(mm * M.nchunks);
      M.Xc64.alloc(mm);
    }
    M.prog.alloc(1);
    M.hprog.alloc(1);
    for (DevBuf<double>* b : {&M.logd_mm, &M.logd_B, &M.b, &M.w, &M.v, &M.tv, &M.Ar}) {
      b->alloc(M.mpad);
      GPRX_HIP_CHECK(hipMemsetAsync(b->p, 0, sizeof(double) * M.mpad, s));
    }
    M.r.alloc(M.npad);
    M.ur.alloc(M.npad);
    const long gp = std::max<long>((M.npad + 255) / 256 * M.mpad, (M.mpad + 255) / 256 * M.mpad);
    M.gpart.alloc(gp);
    M.partials.alloc((size_t)(M.mpad / kTile) * (M.npad / kTile) * kMaxAcc);
    M.partials2.alloc((size_t)cross_tiles(M.mpad, M.mpad) * kMaxAcc);
    M.accs.alloc(3 * kMaxAcc);
    M.scal.alloc(16);
    M.info.alloc(1);
    M.hscal.alloc(16);
    M.hacc.alloc(3 * kMaxAcc);
    M.hinfo.alloc(1);
    GPRX_HIP_CHECK(hipMemsetAsync(M.scal.p, 0, sizeof(double) * 16, s));
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    *out = mo.release();
    return GPRX_OK;
  });
}

int gprx_hip_sgpr_value_and_grad(gprx_hip_model* model, const double* log_theta, size_t p,
                                 double* value, double* grad, gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kSgpr || !log_theta || !value) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle or null output");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* M = static_cast<SgprModel*>(model);
    int rc = exact_validate_theta(M->hp, log_theta, p, err);
    if (rc != GPRX_OK) return rc;
    std::vector<double> g(p, 0.0);
    double v = 0.0;
    rc = M->evaluate(log_theta, grad != nullptr, &v, g.data(), err);
    if (rc != GPRX_OK) return rc;
    M->theta.assign(log_theta, log_theta + p);
    *value = v;
    if (grad) std::memcpy(grad, g.data(), sizeof(double) * p);
    return GPRX_OK;
  });
}

int gprx_hip_sgpr_predict(gprx_hip_model* model, const double* xs, size_t q, size_t d,
                          int32_t variance_kind, double* mean, double* var, gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kSgpr || !mean) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle or null output");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* M = static_cast<SgprModel*>(model);
    if (d != (size_t)M->d) {
      set_error(err, GPRX_LENGTH_MISMATCH, "query dimension differs from training");
      return GPRX_LENGTH_MISMATCH;
    }
    if (q == 0) return GPRX_OK;
    if (!xs || !all_finite(xs, q * d)) {
      set_error(err, GPRX_NON_FINITE_INPUT, "non-finite query");
      return GPRX_NON_FINITE_INPUT;
    }
    return M->predict(xs, q, variance_kind, mean, var, err);
  });
}

}  // extern "C"
