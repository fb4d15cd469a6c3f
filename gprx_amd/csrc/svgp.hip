// Minibatch SVGP (whitened q(u) = N(mu, Lq Lq^T)) on one GPU.
//
// One step follows src/svgp/factor/gradient.rs:104-470 with the explicit
// inverse X = L^-1 of K_mm in place of its triangular solves:
//   A   = X K(Z, X_b)              U = Lq^T A
//   resid_c = y_c - mu'a_c,  var_c = k_cc - |a_c|^2 + |u_c|^2
//   -ELBO = KL - (n/b) sum_c ell_c,  KL = 1/2 (tr S + |mu|^2 - m - log|S|)
//   g_mu   = mu - (n/b)/s2 A resid
//   g_Lq   = [Lq_ij, Lq_ii - 1/Lq_ii] + (n/b)/s2 (A U^T)_ij      (i >= j)
//   G      = (mu resid^T + A - Lq U) / s2
//   w_mn   = X^T G,  w_mm = -sym(X^T tril_half(G A^T) X)
//   g_theta = -(n/b) (<w_mm, dK_mm> + <w_mn, dK(Z, X_b)> - 1/(2 s2) sum dk_cc)
//   g_log s2 = -(n/b) (-b/2 + 1/(2 s2) sum_c (resid_c^2 + var_c))
// The b-long products run on MFMA in the storage type T (f64, or f32 in the
// mixed form); K_mm, its factor and the m x m algebra stay f64.
// Adam (src/optimizer/adam.rs:179-198) runs on the device over the
// unconstrained coordinates of src/svgp/factor/adam.rs:12-83, with batches
// from the reference's epoch shuffle (src/svgp/factor/adam.rs:85-118).
#include <cstring>
#include <memory>
#include <type_traits>

#include "model.hpp"
#include "rng.hpp"

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

// ---------------------------------------------------------------------------
// Step kernels.
// ---------------------------------------------------------------------------
// xb[c + q * bpad] = x[idx[c] + q * ld]; yb[c] = y[idx[c]]; zero beyond b.
__global__ void k_gather(const double* x, long ld, int dp, const double* y, const uint32_t* idx,
                         int b, int bpad, double* xb, double* yb) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= bpad) return;
  if (c < b) {
    const long row = idx[c];
    for (int q = 0; q < dp; ++q) xb[c + (long)q * bpad] = x[row + q * ld];
    yb[c] = y[row];
  } else {
    for (int q = 0; q < dp; ++q) xb[c + (long)q * bpad] = 0.0;
    yb[c] = 0.0;
  }
}

// Lq (ld x ld, lower, zero elsewhere) from the packed columns of params.
template <typename T>
__global__ void k_unpack_lq(const double* packed, int m, long ld, T* L, double* Ld) {
  const long total = ld * ld;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % ld), j = (int)(e / ld);
    double v = 0.0;
    if (i < m && j < m && i >= j) v = packed[(long)j * m - (long)j * (j - 1) / 2 + (i - j)];
    L[e] = (T)v;
    if (Ld) Ld[e] = v;
  }
}

// resid_c, var_c and the one-block sums [0] sum (resid^2 + var).
__global__ __launch_bounds__(1024) void k_point_terms(const double* yb, const double* mu_a,
                                                      const double* a2, const double* u2,
                                                      const double* kxx_p, int b, double* resid,
                                                      double* out) {
  __shared__ double red[16];
  const double kxx = *kxx_p;
  double acc = 0.0;
  for (int c = threadIdx.x; c < b; c += 1024) {
    const double r = yb[c] - mu_a[c];
    const double v = kxx - a2[c] + u2[c];
    resid[c] = r;
    acc += r * r + v;
  }
  for (int c = b + threadIdx.x; c < ((b + 127) / 128) * 128; c += 1024) resid[c] = 0.0;
  acc = block_sum<1024>(acc, red);
  if (threadIdx.x == 0) out[0] = acc;
}

// G = (mu resid^T + A - LU) / s2 over m x b, zero in the padding.
template <typename T>
__global__ void k_adjoint_g(const double* mu, const double* resid, const T* A, const T* LU, long ld,
                            int m, int b, long bpad, double inv_s2, T* G) {
  const long total = ld * bpad;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % ld), c = (int)(e / ld);
    double v = 0.0;
    if (i < m && c < b) v = (mu[i] * resid[c] + (double)A[e] - (double)LU[e]) * inv_s2;
    G[e] = (T)v;
  }
}

// H (f64) = tril_half(W) from storage T.
template <typename T>
__global__ void k_tril_half(const T* W, long ld, double* H) {
  const long total = ld * ld;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % ld), j = (int)(e / ld);
    const double v = (double)W[e];
    H[e] = i > j ? v : (i == j ? 0.5 * v : 0.0);
  }
}

// W = -1/2 (T2 + T2^T)
__global__ void k_neg_sym(const double* T2, long ld, int n, double* W) {
  const long total = (long)n * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    W[i + j * ld] = -0.5 * (T2[i + j * ld] + T2[j + i * ld]);
  }
}

// Variational-parameter gradient (user space) and the KL terms:
// g_mu = mu - c Ares; g_L = [L_ij or L_ii - 1/L_ii] + c M_ij.  Workgroup j
// takes column j of the packed lower triangle (rows i >= j, contiguous in
// both the packed vector and M) and mu_j; part[3 j + {0, 1, 2}] receive its
// sum of L^2, log L_jj and mu_j^2, summed by k_q_sums.
template <typename T>
__global__ __launch_bounds__(256) void k_q_grad(const double* mu, const double* packed, int m,
                                                const double* Ares, const T* Mq, long ld, double c,
                                                double* gmu, double* gL, double* part) {
  __shared__ double red[4];
  const int j = blockIdx.x;
  const long col = (long)j * m - (long)j * (j - 1) / 2;  // packed index of (j, j)
  double s0 = 0.0;
  for (int i = j + (int)threadIdx.x; i < m; i += 256) {
    const long k = col + (i - j);
    const double v = packed[k];
    s0 += v * v;
    const double g = i == j ? v - 1.0 / v : v;
    gL[k] = g + c * (double)Mq[i + (long)j * ld];
  }
  s0 = block_sum<256>(s0, red);
  if (threadIdx.x == 0) {
    part[3 * j] = s0;
    part[3 * j + 1] = log(packed[col]);
    part[3 * j + 2] = mu[j] * mu[j];
    gmu[j] = mu[j] - c * Ares[j];
  }
}

// out[q] = sum_j part[3 j + q], q = 0, 1, 2 (one workgroup, fixed order).
__global__ __launch_bounds__(256) void k_q_sums(const double* part, int m, double* out) {
  __shared__ double red[3][4];
  double s[3] = {0.0, 0.0, 0.0};
  for (int j = threadIdx.x; j < m; j += 256)
    for (int q = 0; q < 3; ++q) s[q] += part[3 * j + q];
  for (int q = 0; q < 3; ++q) s[q] = block_sum<256>(s[q], red[q]);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) out[q] = s[q];
}

// Adam over the variational block: z holds mu (m) then the packed Lq in the
// unconstrained coordinates (diagonal entries in log); g is the user-space
// gradient and user the user-space values.  (src/optimizer/adam.rs:179-198)
__global__ void k_adam_q(double* z, const double* gmu, const double* gL, const double* user_L,
                         double* m1, double* m2, long m, long plen, double lr, double b1, double b2,
                         double eps, double c1, double c2) {
  const long total = m + plen;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    double g;
    if (e < m) {
      g = gmu[e];
    } else {
      const long k = e - m;
      long j = (long)((2.0 * m + 1.0 - sqrt((2.0 * m + 1.0) * (2.0 * m + 1.0) - 8.0 * (double)k)) * 0.5);
      while (j > 0 && j * m - j * (j - 1) / 2 > k) --j;
      while ((j + 1) * m - (j + 1) * j / 2 <= k) ++j;
      const bool diag = k == j * m - j * (j - 1) / 2;
      g = gL[k];
      if (diag) g *= user_L[k];  // d/dlog L_ii
    }
    m1[e] = b1 * m1[e] + (1.0 - b1) * g;
    m2[e] = b2 * m2[e] + (1.0 - b2) * g * g;
    const double mh = m1[e] / c1, vh = m2[e] / c2;
    z[e] -= lr * mh / (sqrt(vh) + eps);
  }
}

// user = map(z): mu as is, packed L with exp on the diagonal.
__global__ void k_q_to_user(const double* z, double* user, long m, long plen) {
  const long total = m + plen;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    double v = z[e];
    if (e >= m) {
      const long k = e - m;
      long j = (long)((2.0 * m + 1.0 - sqrt((2.0 * m + 1.0) * (2.0 * m + 1.0) - 8.0 * (double)k)) * 0.5);
      while (j > 0 && j * m - j * (j - 1) / 2 > k) --j;
      while ((j + 1) * m - (j + 1) * j / 2 <= k) ++j;
      if (k == j * m - j * (j - 1) / 2) v = exp(v);
    }
    user[e] = v;
  }
}

__global__ void k_q_to_z(const double* user, double* z, long m, long plen) {
  const long total = m + plen;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    double v = user[e];
    if (e >= m) {
      const long k = e - m;
      long j = (long)((2.0 * m + 1.0 - sqrt((2.0 * m + 1.0) * (2.0 * m + 1.0) - 8.0 * (double)k)) * 0.5);
      while (j > 0 && j * m - j * (j - 1) / 2 > k) --j;
      while ((j + 1) * m - (j + 1) * j / 2 <= k) ++j;
      if (k == j * m - j * (j - 1) / 2) v = log(v);
    }
    z[e] = v;
  }
}

inline int grid_for(long total) { return (int)std::min<long>((total + 255) / 256, 65536); }

struct SvgpModel : gprx_hip_model {
  SvgpModel() : gprx_hip_model(kSvgp) {}
  ~SvgpModel() override {
    if (stream) (void)hipStreamSynchronize(stream);
    fstreams.destroy();
    if (stream) {
      for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
      for (auto& e : sev)
        if (e) (void)hipEventDestroy(e);
      (void)hipStreamDestroy(stream);
    }
  }
  HostProgram hp;
  int n = 0, m = 0, d = 0, dp = 0, mt = 0, n_theta = 0;
  long npad = 0, mpad = 0, ldt = 0, bcap = 0;
  size_t plen = 0, ptot = 0;
  int precision = GPRX_FP64;
  hipStream_t stream = nullptr;
  FactorStreams fstreams;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // sev[0]: the step's program and batch indices are on the device; sev[1]:
  // the factor-independent start of the step (on aux) is done
  hipEvent_t sev[2] = {nullptr, nullptr};
  // the factor's lookahead stream (shared, not owned): the step's
  // factor-independent start runs there, so the process holds no further
  // CU-masked queue
  hipStream_t aux = nullptr;
  DevBuf<double> x, y, z;
  DevBuf<DevProgram> prog;
  HostBuf<DevProgram> hprog;
  // f64 m x m
  DevBuf<double> KA, Xmm, Tm, Ld, H, T1, T2, Wmm, Mq64, Wpre64;
  // storage-typed
  DevBuf<double> X64c, L64c, A64, U64, G64, LU64, Wmn64, Kzx64;
  DevBuf<float> X32, L32, A32, U32, G32, LU32, Wmn32, Kzx32, Mq32, Wpre32;
  DevBuf<double> xb, yb, mu_a, a2, u2, resid, Ares, logd, scal, partials, partials2, accs, gmu, gL, qpart;
  DevBuf<double> params_d;  // user-space params (device)
  DevBuf<double> zq, m1q, m2q, userq;  // Adam state of the variational block
  DevBuf<uint32_t> perm, bidx;
  DevBuf<int> info;
  HostBuf<double> hscal, hacc;
  HostBuf<int> hinfo;

  bool mixed() const { return precision == GPRX_MIXED; }
  DevPoints zp() const { return DevPoints{z.p, mpad, m}; }

  void upload_program(const double* th, double jitter) {
    make_dev_program(hp, th, jitter, hprog.p);
    GPRX_HIP_CHECK(hipMemcpyAsync(prog.p, hprog.p, sizeof(DevProgram), hipMemcpyHostToDevice, stream));
  }

  void ensure_batch(long b) {
    if (b <= bcap) return;
    bcap = (long)round_up((size_t)b, kTile);
    const size_t mb = (size_t)mpad * bcap;
    xb.alloc((size_t)bcap * dp);
    yb.alloc(bcap);
    mu_a.alloc(bcap);
    a2.alloc(bcap);
    u2.alloc(bcap);
    resid.alloc(bcap);
    bidx.alloc(bcap);
    if (mixed()) {
      for (DevBuf<float>* p : {&A32, &U32, &G32, &LU32, &Wmn32, &Kzx32}) p->alloc(mb);
    } else {
      for (DevBuf<double>* p : {&A64, &U64, &G64, &LU64, &Wmn64, &Kzx64}) p->alloc(mb);
    }
    partials2.alloc((size_t)(mpad / kTile) * (bcap / kTile) * kMaxAcc);
  }

  // Fill and factor K_mm at jitter j (info bits in info; no wait).
  void launch_factor(const double* th, double j) {
    upload_program(th, j);
    GPRX_HIP_CHECK(hipEventRecord(sev[0], stream));
    GPRX_HIP_CHECK(hipMemsetAsync(info.p, 0, sizeof(int), stream));
    fill_lower<double>(dp, prog.p, zp(), KA.p, mpad, nullptr, info.p, stream);
    factor_inverse<double>(KA.p, Xmm.p, Tm.p, mpad, ldt, mt, logd.p, info.p, fstreams);
  }

  // K_mm factor with the adaptive default jitter; returns a status.
  int factor_kmm(const double* th, gprx_error* err) {
    double j = 0.0, next = 1e-8;
    int remaining = 5;
    for (;;) {
      launch_factor(th, j);
      GPRX_HIP_CHECK(hipMemcpyAsync(hinfo.p, info.p, sizeof(int), hipMemcpyDeviceToHost, stream));
      GPRX_HIP_CHECK(hipStreamSynchronize(stream));
      const int bits = *hinfo.p;
      if (bits == 0) return GPRX_OK;
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
  }

  // One step at the user parameters held on the device (params_d; the
  // kernel + noise part also in th on the host) over the batch in bidx
  // (b indices).  Writes value, the theta gradient (host, n_theta) and the
  // variational gradient (device gmu, gL).
  // The K_mm factor's first try (no jitter) is checked only with the
  // step's own results, so the common case runs the whole step without a
  // host round trip; if it failed, the jitter loop runs and the step is
  // evaluated again.
  template <typename T>
  int step(const double* th, long b, double* value, double* gtheta, gprx_error* err,
           T* Xc, T* Lc, T* Kzx, T* A, T* U, T* G, T* LU, T* Wmn, T* Mq, T* Wpre) {
    launch_factor(th, 0.0);
    const int rc = step_body<T>(th, b, value, gtheta, err, Xc, Lc, Kzx, A, U, G, LU, Wmn, Mq, Wpre);
    if (*hinfo.p == 0) return rc;
    const int frc = factor_kmm(th, err);
    if (frc != GPRX_OK) return frc;
    return step_body<T>(th, b, value, gtheta, err, Xc, Lc, Kzx, A, U, G, LU, Wmn, Mq, Wpre);
  }

  // The step after K_mm's factor has been launched; leaves the factor's
  // info bits in hinfo.
  template <typename T>
  int step_body(const double* th, long b, double* value, double* gtheta, gprx_error* err,
                T* Xc, T* Lc, T* Kzx, T* A, T* U, T* G, T* LU, T* Wmn, T* Mq, T* Wpre) {
    hipStream_t s = stream;
    const int P = hp.n_params;
    const double s2 = std::exp(th[P]);
    const double scale = (double)n / (double)b;
    GPRX_HIP_CHECK(hipEventRecord(ev[0], s));
    const long bp = (long)round_up((size_t)b, kTile);
    const int M = (int)mpad, B = (int)bp;
    const double* pu = params_d.p;
    DevPoints xbp{xb.p, bp, (int)b};
    // What does not depend on K_mm's factor runs on aux beside it, from the
    // moment the program and the batch indices are on the device (sev[0],
    // recorded by launch_factor): the batch's points, K(Z, X_b), Lq and
    // k(x, x).
    GPRX_HIP_CHECK(hipStreamWaitEvent(aux, sev[0], 0));
    hipLaunchKernelGGL(k_gather, dim3((B + 255) / 256), dim3(256), 0, aux, x.p, npad, dp, y.p, bidx.p,
                       (int)b, B, xb.p, yb.p);
    hipLaunchKernelGGL((k_unpack_lq<T>), dim3(grid_for(mpad * mpad)), dim3(256), 0, aux,
                       pu + n_theta + m, m, mpad, Lc, Ld.p);
    fill_cross<T>(dp, prog.p, zp(), xbp, Kzx, mpad, aux);
    fill_diag(dp, prog.p, DevPoints{z.p, mpad, 1}, scal.p + 7, aux);
    GPRX_HIP_CHECK(hipEventRecord(sev[1], aux));
    // storage copy of X = L^-1
    if (std::is_same<T, double>::value) {
      GPRX_HIP_CHECK(hipMemcpyAsync(Xc, Xmm.p, sizeof(double) * mpad * mpad, hipMemcpyDeviceToDevice, s));
    } else {
      convert<float, double>(reinterpret_cast<float*>(Xc), Xmm.p, (size_t)mpad * mpad, s);
    }
    GPRX_HIP_CHECK(hipStreamWaitEvent(s, sev[1], 0));
    // A = X Kzx; U = Lq^T A
    gemm_call<T>(false, false, Xc, mpad, Kzx, mpad, A, mpad, M, B, M, 1.0, 0.0, kKhiRow, 0,
                 kOrderColFastRev, s);
    gemm_call<T>(true, false, Lc, mpad, A, mpad, U, mpad, M, B, M, 1.0, 0.0, kKloRow, 0,
                 kOrderColFast, s);
    gemv_t<T>(A, mpad, M, B, pu + n_theta, mu_a.p, 0, 0, 0, 0, s);
    col_sumsq<T>(A, mpad, M, B, a2.p, s);
    col_sumsq<T>(U, mpad, M, B, u2.p, s);
    // resid, var; scal[0] = sum (resid^2 + var)  (k(x,x) read from scal[7])
    hipLaunchKernelGGL(k_point_terms, dim3(1), dim3(1024), 0, s, yb.p, mu_a.p, a2.p, u2.p,
                       scal.p + 7, (int)b, resid.p, scal.p);
    gemv_n<T>(A, mpad, M, B, resid.p, Ares.p, partials.p, 0, 0, 0, s);
    // Mq = A U^T (lower used)
    gemm_call<T>(false, true, A, mpad, U, mpad, Mq, mpad, M, M, B, 1.0, 0.0, 0, 1, kOrderRowFast, s);
    hipLaunchKernelGGL((k_q_grad<T>), dim3(m), dim3(256), 0, s, pu + n_theta, pu + n_theta + m, m,
                       Ares.p, Mq, mpad, scale / s2, gmu.p, gL.p, qpart.p);
    hipLaunchKernelGGL(k_q_sums, dim3(1), dim3(256), 0, s, qpart.p, m, scal.p + 1);
    // G = (mu resid^T + A - Lq U) / s2
    gemm_call<T>(false, false, Lc, mpad, U, mpad, LU, mpad, M, B, M, 1.0, 0.0, kKhiRow, 0,
                 kOrderColFastRev, s);
    hipLaunchKernelGGL((k_adjoint_g<T>), dim3(grid_for(mpad * bp)), dim3(256), 0, s, pu + n_theta,
                       resid.p, A, LU, mpad, m, (int)b, bp, 1.0 / s2, G);
    // Wpre = G A^T; w_mn = X^T G
    gemm_call<T>(false, true, G, mpad, A, mpad, Wpre, mpad, M, M, B, 1.0, 0.0, 0, 0, kOrderRowFast, s);
    gemm_call<T>(true, false, Xc, mpad, G, mpad, Wmn, mpad, M, B, M, 1.0, 0.0, kKloRow, 0,
                 kOrderColFast, s);
    // w_mm = -sym(X^T tril_half(Wpre) X)
    hipLaunchKernelGGL((k_tril_half<T>), dim3(grid_for(mpad * mpad)), dim3(256), 0, s, Wpre, mpad, H.p);
    gemm_call<double>(true, false, Xmm.p, mpad, H.p, mpad, T1.p, mpad, M, M, M, 1.0, 0.0, kKloRow, 0,
                      kOrderRowFast, s);
    gemm_call<double>(false, false, T1.p, mpad, Xmm.p, mpad, T2.p, mpad, M, M, M, 1.0, 0.0, kKloCol,
                      0, kOrderRowFast, s);
    hipLaunchKernelGGL(k_neg_sym, dim3(grid_for(mpad * mpad)), dim3(256), 0, s, T2.p, mpad, (int)mpad,
                       Wmm.p);
    contract_cross<double>(dp, hp.ns, prog.p, zp(), zp(), Wmm.p, mpad, partials.p, s);
    reduce_partials(partials.p, cross_tiles(mpad, mpad), accs.p, s);
    contract_cross<T>(dp, hp.ns, prog.p, zp(), xbp, Wmn, mpad, partials2.p, s);
    reduce_partials(partials2.p, cross_tiles(mpad, bp), accs.p + kMaxAcc, s);
    contract_diag(dp, prog.p, xbp, -0.5 / s2, accs.p + 2 * kMaxAcc, s);
    GPRX_HIP_CHECK(hipEventRecord(ev[1], s));
    GPRX_HIP_CHECK(hipMemcpyAsync(hscal.p, scal.p, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(hacc.p, accs.p, sizeof(double) * 3 * kMaxAcc, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(hinfo.p, info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    if (*hinfo.p) return GPRX_OK;  // the caller factors again with jitter
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
    last_eval_ms = ms;
    const double r2v = hscal.p[0], sumL2 = hscal.p[1], sumlogL = hscal.p[2], mu2 = hscal.p[3];
    const double kl = 0.5 * (sumL2 + mu2 - (double)m - 2.0 * sumlogL);
    const double ell = -0.5 * (double)b * (kLog2Pi + std::log(s2)) - 0.5 * r2v / s2;
    *value = kl - scale * ell;
    if (!std::isfinite(*value)) {
      set_error(err, GPRX_NON_FINITE_KERNEL_VALUE, "non-finite -ELBO");
      return GPRX_NON_FINITE_KERNEL_VALUE;
    }
    double acc[kMaxAcc];
    for (int k = 0; k < kMaxAcc; ++k) acc[k] = hacc.p[k] + hacc.p[kMaxAcc + k] + hacc.p[2 * kMaxAcc + k];
    acc_to_grad(hp, th, acc, gtheta);
    for (int p = 0; p < P; ++p) gtheta[p] *= -scale;
    gtheta[P] = -scale * (-0.5 * (double)b + 0.5 * r2v / s2);
    return GPRX_OK;
  }

  int run_step(const double* th, long b, double* value, double* gtheta, gprx_error* err) {
    ensure_batch(b);
    if (mixed())
      return step<float>(th, b, value, gtheta, err, X32.p, L32.p, Kzx32.p, A32.p, U32.p, G32.p,
                         LU32.p, Wmn32.p, Mq32.p, Wpre32.p);
    return step<double>(th, b, value, gtheta, err, X64c.p, L64c.p, Kzx64.p, A64.p, U64.p, G64.p,
                        LU64.p, Wmn64.p, Mq64.p, Wpre64.p);
  }
};

}  // namespace

extern "C" {

int gprx_hip_svgp_create(const double* x, size_t n, size_t d, const double* y, const double* z,
                         size_t m, const gprx_kernel_program* kernel, const double* log_theta,
                         size_t n_theta, int32_t precision, gprx_hip_model** out, gprx_error* err) {
  if (out) *out = nullptr;
  return guarded(err, [&]() -> int {
    if (!x || !y || !z || !out || n == 0 || d == 0 || m == 0) {
      set_error(err, GPRX_EMPTY_INPUT, "empty data or inducing set");
      return GPRX_EMPTY_INPUT;
    }
    if (d > (size_t)kMaxDim || m > 8192 || n > ((size_t)1 << 28)) {
      set_error(err, GPRX_UNSUPPORTED, "svgp sizes above d = 32, m = 8192 or n = 2^28");
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
    auto mo = std::make_unique<SvgpModel>();
    SvgpModel& M = *mo;
    std::string why;
    int rc = parse_program(kernel, (int)d, &M.hp, &why);
    if (rc != GPRX_OK) {
      set_error(err, rc, why);
      return rc;
    }
    if (!log_theta) {
      set_error(err, GPRX_LENGTH_MISMATCH, "missing parameters");
      return GPRX_LENGTH_MISMATCH;
    }
    rc = exact_validate_theta(M.hp, log_theta, n_theta, err);
    if (rc != GPRX_OK) return rc;
    M.n = (int)n;
    M.m = (int)m;
    M.d = (int)d;
    M.dp = padded_dim((int)d);
    M.n_theta = (int)n_theta;
    M.npad = (long)round_up(n, kTile);
    M.mpad = (long)round_up(m, kTile);
    M.mt = (int)(M.mpad / kTile);
    M.ldt = (long)((M.mt + 1) / 2) * kTile;
    M.plen = m * (m + 1) / 2;
    M.ptot = n_theta + m + M.plen;
    M.precision = precision;
    GPRX_HIP_CHECK(hipStreamCreateWithFlags(&M.stream, hipStreamNonBlocking));
    M.fstreams.init(M.stream);
    for (auto& e : M.ev) GPRX_HIP_CHECK(hipEventCreate(&e));
    for (auto& e : M.sev) GPRX_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    M.aux = ahead_stream();
    hipStream_t s = M.stream;
    exact_upload_points(M.x, x, (int)n, (int)d, M.npad, M.dp, s);
    exact_upload_points(M.z, z, (int)m, (int)d, M.mpad, M.dp, s);
    M.y.alloc(M.npad);
    GPRX_HIP_CHECK(hipMemsetAsync(M.y.p, 0, sizeof(double) * M.npad, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(M.y.p, y, sizeof(double) * n, hipMemcpyHostToDevice, s));
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    const size_t mm = (size_t)M.mpad * M.mpad;
    for (DevBuf<double>* p : {&M.KA, &M.Xmm, &M.Ld, &M.H, &M.T1, &M.T2, &M.Wmm}) {
      p->alloc(mm);
      GPRX_HIP_CHECK(hipMemsetAsync(p->p, 0, sizeof(double) * mm, s));
    }
    M.Tm.alloc((size_t)M.ldt * factor_tmp_cols(M.mt) * kTile);
    if (precision == GPRX_MIXED) {
      for (DevBuf<float>* p : {&M.X32, &M.L32, &M.Mq32, &M.Wpre32}) p->alloc(mm);
    } else {
      for (DevBuf<double>* p : {&M.X64c, &M.L64c, &M.Mq64, &M.Wpre64}) p->alloc(mm);
    }
    M.prog.alloc(1);
    M.hprog.alloc(1);
    M.logd.alloc(M.mpad);
    M.Ares.alloc(M.mpad);
    M.gmu.alloc(m);
    M.gL.alloc(M.plen);
    M.qpart.alloc(3 * (size_t)m);
    M.params_d.alloc(M.ptot);
    M.scal.alloc(16);
    M.accs.alloc(3 * kMaxAcc);
    const long gp = std::max<long>((long)cross_tiles(M.mpad, M.mpad) * kMaxAcc, 64L * M.mpad);
    M.partials.alloc(gp);
    M.info.alloc(1);
    M.hscal.alloc(16);
    M.hacc.alloc(3 * kMaxAcc);
    M.hinfo.alloc(1);
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    *out = mo.release();
    return GPRX_OK;
  });
}

size_t gprx_hip_svgp_num_params(const gprx_hip_model* model) {
  if (!model || model->kind != gprx_hip_model::kSvgp) return 0;
  return static_cast<const SvgpModel*>(model)->ptot;
}

int gprx_hip_svgp_value_and_grad(gprx_hip_model* model, const double* params, size_t p,
                                 const uint64_t* batch, size_t b, double* value, double* grad,
                                 gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kSvgp || !params || !value) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle or null output");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* M = static_cast<SvgpModel*>(model);
    if (p != M->ptot) {
      set_error(err, GPRX_LENGTH_MISMATCH, "parameter count differs");
      return GPRX_LENGTH_MISMATCH;
    }
    if (b == 0 || !batch) {
      set_error(err, GPRX_EMPTY_INPUT, "empty batch");
      return GPRX_EMPTY_INPUT;
    }
    if (!all_finite(params, p)) {
      set_error(err, GPRX_NON_FINITE_INPUT, "non-finite parameters");
      return GPRX_NON_FINITE_INPUT;
    }
    int rc = exact_validate_theta(M->hp, params, M->n_theta, err);
    if (rc != GPRX_OK) return rc;
    for (size_t j = 0, k = 0; j < (size_t)M->m; ++j)
      for (size_t i = j; i < (size_t)M->m; ++i, ++k)
        if (i == j && !(params[M->n_theta + M->m + k] > 0.0)) {
          set_error(err, GPRX_INVALID_HYPERPARAMETER, "variational Cholesky diagonal must be positive");
          return GPRX_INVALID_HYPERPARAMETER;
        }
    std::vector<uint32_t> idx(b);
    for (size_t c = 0; c < b; ++c) {
      if (batch[c] >= (uint64_t)M->n) {
        set_error(err, GPRX_LENGTH_MISMATCH, "batch index out of range");
        return GPRX_LENGTH_MISMATCH;
      }
      idx[c] = (uint32_t)batch[c];
    }
    M->ensure_batch((long)b);
    GPRX_HIP_CHECK(hipMemcpyAsync(M->params_d.p, params, sizeof(double) * p, hipMemcpyHostToDevice, M->stream));
    GPRX_HIP_CHECK(hipMemcpyAsync(M->bidx.p, idx.data(), sizeof(uint32_t) * b, hipMemcpyHostToDevice, M->stream));
    std::vector<double> gth(M->n_theta);
    double v = 0.0;
    rc = M->run_step(params, (long)b, &v, gth.data(), err);
    if (rc != GPRX_OK) return rc;
    *value = v;
    if (grad) {
      std::memcpy(grad, gth.data(), sizeof(double) * M->n_theta);
      GPRX_HIP_CHECK(hipMemcpyAsync(grad + M->n_theta, M->gmu.p, sizeof(double) * M->m, hipMemcpyDeviceToHost, M->stream));
      GPRX_HIP_CHECK(hipMemcpyAsync(grad + M->n_theta + M->m, M->gL.p, sizeof(double) * M->plen, hipMemcpyDeviceToHost, M->stream));
      GPRX_HIP_CHECK(hipStreamSynchronize(M->stream));
    }
    return GPRX_OK;
  });
}

int gprx_hip_svgp_adam(gprx_hip_model* model, double* params, size_t p, size_t batch_size,
                       size_t steps, double lr, double beta1, double beta2, double eps,
                       uint64_t seed, double* values, gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kSvgp || !params) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle or null output");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* M = static_cast<SvgpModel*>(model);
    if (p != M->ptot || batch_size == 0) {
      set_error(err, GPRX_LENGTH_MISMATCH, "parameter count or batch size");
      return GPRX_LENGTH_MISMATCH;
    }
    int rc = exact_validate_theta(M->hp, params, M->n_theta, err);
    if (rc != GPRX_OK) return rc;
    hipStream_t s = M->stream;
    const int nth = M->n_theta;
    const long mq = M->m, pl = (long)M->plen;
    const size_t nq = (size_t)(mq + pl);
    // theta through the interval logit of src/optimizer/logit.rs with the
    // default (1e-5, 1e5) interval; the variational block on the device.
    const double ln_lo = std::log(1e-5), ln_hi = std::log(1e5), span = ln_hi - ln_lo, lscale = span / 4.0;
    std::vector<double> zt(nth), m1t(nth, 0.0), m2t(nth, 0.0), th(nth), gth(nth);
    for (int i = 0; i < nth; ++i) {
      double u = (params[i] - ln_lo) / span;
      if (!(u > 0.0 && u < 1.0)) {
        set_error(err, GPRX_INVALID_HYPERPARAMETER, "parameter outside (1e-5, 1e5)");
        return GPRX_INVALID_HYPERPARAMETER;
      }
      u = std::min(std::max(u, 2.220446049250313e-16), 1.0 - 2.220446049250313e-16);
      zt[i] = std::log(u / (1.0 - u)) * lscale;
    }
    M->zq.alloc(nq);
    M->m1q.alloc(nq);
    M->m2q.alloc(nq);
    M->userq.alloc(nq);
    GPRX_HIP_CHECK(hipMemcpyAsync(M->userq.p, params + nth, sizeof(double) * nq, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_q_to_z, dim3(grid_for((long)nq)), dim3(256), 0, s, M->userq.p, M->zq.p, mq, pl);
    GPRX_HIP_CHECK(hipMemsetAsync(M->m1q.p, 0, sizeof(double) * nq, s));
    GPRX_HIP_CHECK(hipMemsetAsync(M->m2q.p, 0, sizeof(double) * nq, s));
    const long n = M->n;
    std::vector<uint32_t> order(n);
    for (long i = 0; i < n; ++i) order[i] = (uint32_t)i;
    M->perm.alloc(n);
    M->ensure_batch((long)std::min<size_t>(batch_size, (size_t)n));
    Rng rng(seed);
    long start = n;  // forces a shuffle first
    uint64_t t = 0;
    auto sigmoid = [](double v) {
      if (v >= 0.0) {
        const double e = std::exp(-v);
        return 1.0 / (1.0 + e);
      }
      const double e = std::exp(v);
      return e / (1.0 + e);
    };
    for (size_t st = 0; st < steps; ++st) {
      if (start >= n) {
        for (long i = n - 1; i >= 1; --i) {
          const long j = (long)rng.up_to((uint64_t)i);
          std::swap(order[i], order[j]);
        }
        GPRX_HIP_CHECK(hipMemcpyAsync(M->perm.p, order.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
        start = 0;
      }
      const long end = std::min<long>(start + (long)batch_size, n);
      const long b = end - start;
      GPRX_HIP_CHECK(hipMemcpyAsync(M->bidx.p, M->perm.p + start, sizeof(uint32_t) * b, hipMemcpyDeviceToDevice, s));
      // user parameters
      for (int i = 0; i < nth; ++i) th[i] = ln_lo + sigmoid(zt[i] / lscale) * span;
      GPRX_HIP_CHECK(hipMemcpyAsync(M->params_d.p, th.data(), sizeof(double) * nth, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_q_to_user, dim3(grid_for((long)nq)), dim3(256), 0, s, M->zq.p,
                         M->params_d.p + nth, mq, pl);
      double v = 0.0;
      rc = M->run_step(th.data(), b, &v, gth.data(), err);
      if (rc != GPRX_OK) return rc;
      if (values) values[st] = v;
      // Adam (src/optimizer/adam.rs:179-198)
      ++t;
      const double c1 = 1.0 - std::pow(beta1, (double)t), c2 = 1.0 - std::pow(beta2, (double)t);
      for (int i = 0; i < nth; ++i) {
        const double sg = sigmoid(zt[i] / lscale);
        const double g = gth[i] * span * sg * (1.0 - sg) / lscale;
        m1t[i] = beta1 * m1t[i] + (1.0 - beta1) * g;
        m2t[i] = beta2 * m2t[i] + (1.0 - beta2) * g * g;
        zt[i] -= lr * (m1t[i] / c1) / (std::sqrt(m2t[i] / c2) + eps);
      }
      hipLaunchKernelGGL(k_adam_q, dim3(grid_for((long)nq)), dim3(256), 0, s, M->zq.p + 0, M->gmu.p,
                         M->gL.p, M->params_d.p + nth + mq, M->m1q.p, M->m2q.p, mq, pl, lr, beta1,
                         beta2, eps, c1, c2);
      start = end;
    }
    for (int i = 0; i < nth; ++i) params[i] = ln_lo + sigmoid(zt[i] / lscale) * span;
    hipLaunchKernelGGL(k_q_to_user, dim3(grid_for((long)nq)), dim3(256), 0, s, M->zq.p, M->userq.p, mq, pl);
    GPRX_HIP_CHECK(hipMemcpyAsync(params + nth, M->userq.p, sizeof(double) * nq, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    return GPRX_OK;
  });
}

int gprx_hip_svgp_predict(gprx_hip_model* model, const double* params, size_t p, const double* xs,
                          size_t q, size_t d, int32_t variance_kind, double* mean, double* var,
                          gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kSvgp || !params || !mean) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle or null output");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* M = static_cast<SvgpModel*>(model);
    if (p != M->ptot || d != (size_t)M->d) {
      set_error(err, GPRX_LENGTH_MISMATCH, "parameter count or dimension");
      return GPRX_LENGTH_MISMATCH;
    }
    if (q == 0) return GPRX_OK;
    if (!xs || !all_finite(xs, q * d)) {
      set_error(err, GPRX_NON_FINITE_INPUT, "non-finite query");
      return GPRX_NON_FINITE_INPUT;
    }
    int rc = M->factor_kmm(params, err);
    if (rc != GPRX_OK) {
      if (err) err->stage = GPRX_STAGE_PREDICT;
      return rc;
    }
    hipStream_t s = M->stream;
    const long qpad = (long)round_up(q, kTile);
    const int Mi = (int)M->mpad, Q = (int)qpad;
    DevBuf<double> xq, Ks, A, U, Lq, mu_d, mu, a2, u2, kss;
    exact_upload_points(xq, xs, (int)q, (int)d, qpad, M->dp, s);
    DevPoints pq{xq.p, qpad, (int)q};
    Ks.alloc((size_t)M->mpad * qpad);
    A.alloc((size_t)M->mpad * qpad);
    U.alloc((size_t)M->mpad * qpad);
    Lq.alloc((size_t)M->mpad * M->mpad);
    mu_d.alloc(M->mpad);
    mu.alloc(qpad);
    a2.alloc(qpad);
    u2.alloc(qpad);
    kss.alloc(qpad);
    GPRX_HIP_CHECK(hipMemsetAsync(mu_d.p, 0, sizeof(double) * M->mpad, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(mu_d.p, params + M->n_theta, sizeof(double) * M->m, hipMemcpyHostToDevice, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(M->params_d.p, params, sizeof(double) * p, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_unpack_lq<double>), dim3(grid_for(M->mpad * M->mpad)), dim3(256), 0, s,
                       M->params_d.p + M->n_theta + M->m, M->m, M->mpad, Lq.p, (double*)nullptr);
    fill_cross<double>(M->dp, M->prog.p, M->zp(), pq, Ks.p, M->mpad, s);
    gemm_call<double>(false, false, M->Xmm.p, M->mpad, Ks.p, M->mpad, A.p, M->mpad, Mi, Q, Mi, 1.0,
                      0.0, kKhiRow, 0, kOrderColFastRev, s);
    gemm_call<double>(true, false, Lq.p, M->mpad, A.p, M->mpad, U.p, M->mpad, Mi, Q, Mi, 1.0, 0.0,
                      kKloRow, 0, kOrderColFast, s);
    gemv_t<double>(A.p, M->mpad, Mi, Q, mu_d.p, mu.p, 0, 0, 0, 0, s);
    col_sumsq<double>(A.p, M->mpad, Mi, Q, a2.p, s);
    col_sumsq<double>(U.p, M->mpad, Mi, Q, u2.p, s);
    fill_diag(M->dp, M->prog.p, pq, kss.p, s);
    std::vector<double> h1(q), h2(q), hk(q);
    GPRX_HIP_CHECK(hipMemcpyAsync(mean, mu.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(h1.data(), a2.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(h2.data(), u2.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(hk.data(), kss.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    if (var) {
      const double noise = std::exp(params[M->n_theta - 1]);
      for (size_t k = 0; k < q; ++k) {
        double lat = hk[k] - h1[k] + h2[k];
        if (lat < 0.0) lat = 0.0;
        var[k] = variance_kind == GPRX_VARIANCE_LATENT ? lat : lat + noise;
      }
    }
    return GPRX_OK;
  });
}

}  // extern "C"
