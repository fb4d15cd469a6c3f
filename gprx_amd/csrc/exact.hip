// Exact GPR on one GPU: the NLML + gradient evaluation behind gprx's
// objective seam, and the predictive moments.
//
// One evaluation (src/gpr/exact_fit.rs:319-341, 665-729):
//   fill      A = K(x, x) + (sigma^2 + j) I            (lower tiles)
//   factor    L = chol(A), X = L^-1                      (recursive, MFMA)
//   alpha     alpha = X^T (X y)                          (FP64)   or
//             alpha refined with f64 residuals           (MIXED, src/precision/refine.rs:56-98)
//   value     1/2 y'alpha + sum log L_ii + n/2 log 2 pi  (src/gpr/factor.rs:254-267)
//   inverse   A^-1 = X^T X                               (lower tiles, MFMA)
//   gradient  -1/2 <alpha alpha' - A^-1, dK/dtheta>      (fused: K recomputed from x)
//             -1/2 sigma^2 sum_i (alpha_i^2 - A^-1_ii)   (src/gpr/exact_fit.rs:676-682)
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>

#include "hessian.hpp"
#include "model.hpp"

using namespace gprx;

namespace {

constexpr double kLog2Pi = 1.8378770664093453;  // log(2 pi)

struct ExactModel : gprx_hip_model {
  ExactModel() : gprx_hip_model(kExact) {}
  ~ExactModel() override {
    if (stream) (void)hipStreamSynchronize(stream);
    fstreams.destroy();
    if (stream) {
      for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
      (void)hipStreamDestroy(stream);
    }
  }
  HostProgram hp;
  int n = 0, d = 0, dp = 0, nt = 0;
  long npad = 0, ldt = 0;
  int precision = GPRX_FP64;
  gprx_jitter_policy jit{0.0, 1.0, 0.0, 0};
  // Largest jitter the mixed form's f32 factor may take; a factor that needs
  // more is formed in FP64 instead (0: the f32 factor adds none of its own).
  double mixed_jitter_bound = 0.0;
  std::vector<double> theta;
  double last_jitter = 0.0;
  int fallback = 0, refine_steps = 0;  // fallback: GPRX_FALLBACK_* of the last evaluation
  bool factored = false;
  int cur_form = GPRX_FP64;  // which buffers hold X and alpha
  hipStream_t stream = nullptr;
  FactorStreams fstreams;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double phase_ms[5] = {0, 0, 0, 0, 0};
  DevBuf<double> x, y;
  DevBuf<DevProgram> prog;
  HostBuf<DevProgram> hprog;
  DevBuf<double> A64, X64, T64, K64;
  DevBuf<float> A32, X32, T32;
  DevBuf<double> logdiag, alpha, tv, r, dlt, p1, q1, tpart, partials, accs, scal;
  DevBuf<int> info;
  HostBuf<double> hscal, hacc;
  HostBuf<int> hinfo;
  // online insertion: host copies of the data (point-major: point i at
  // hx[i d .. i d + d), so an insertion appends d values), and device work
  // vectors
  std::vector<double> hx, hy;
  DevBuf<double> zv, kcol, lv, wv, xnew;
  bool z_valid = false;  // zv = X y for the factor in X64

  DevPoints pts() const { return DevPoints{x.p, npad, n}; }

  void alloc_fp64() {
    if (A64.p) return;
    A64.alloc(npad * npad);
    X64.alloc(npad * npad);
    T64.alloc(ldt * factor_tmp_cols(nt) * kTile);
    GPRX_HIP_CHECK(hipMemsetAsync(A64.p, 0, sizeof(double) * npad * npad, stream));
    GPRX_HIP_CHECK(hipMemsetAsync(X64.p, 0, sizeof(double) * npad * npad, stream));
  }
  void alloc_mixed() {
    if (A32.p) return;
    A32.alloc(npad * npad);
    X32.alloc(npad * npad);
    T32.alloc(ldt * factor_tmp_cols(nt) * kTile);
    K64.alloc(npad * npad);
    GPRX_HIP_CHECK(hipMemsetAsync(A32.p, 0, sizeof(float) * npad * npad, stream));
    GPRX_HIP_CHECK(hipMemsetAsync(X32.p, 0, sizeof(float) * npad * npad, stream));
    GPRX_HIP_CHECK(hipMemsetAsync(K64.p, 0, sizeof(double) * npad * npad, stream));
  }

  void upload_program(const double* th, double jitter) {
    make_dev_program(hp, th, std::exp(th[hp.n_params]) + jitter, hprog.p);
    GPRX_HIP_CHECK(hipMemcpyAsync(prog.p, hprog.p, sizeof(DevProgram), hipMemcpyHostToDevice, stream));
  }

  // Reads back info and the scalars; returns the info bits.
  int readback(int nscal, bool with_acc) {
    GPRX_HIP_CHECK(hipMemcpyAsync(hinfo.p, info.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    GPRX_HIP_CHECK(hipMemcpyAsync(hscal.p, scal.p, sizeof(double) * nscal, hipMemcpyDeviceToHost, stream));
    if (with_acc)
      GPRX_HIP_CHECK(hipMemcpyAsync(hacc.p, accs.p, sizeof(double) * kMaxAcc, hipMemcpyDeviceToHost, stream));
    GPRX_HIP_CHECK(hipStreamSynchronize(stream));
    return *hinfo.p;
  }

  // FP64 form.  Returns info bits (1: factor failed, 2: non-finite kernel).
  int run_fp64(bool want_grad) {
    alloc_fp64();
    hipStream_t s = stream;
    GPRX_HIP_CHECK(hipMemsetAsync(info.p, 0, sizeof(int), s));
    fill_lower<double>(dp, prog.p, pts(), A64.p, npad, nullptr, info.p, s);
    GPRX_HIP_CHECK(hipEventRecord(ev[1], s));
    factor_inverse<double>(A64.p, X64.p, T64.p, npad, ldt, nt, logdiag.p, info.p, fstreams);
    GPRX_HIP_CHECK(hipEventRecord(ev[2], s));
    tile_mv_lower<double>(X64.p, npad, nt, kTmvRow, y.p, tv.p, nullptr, tpart.p, s);
    tile_mv_lower<double>(X64.p, npad, nt, kTmvCol, tv.p, alpha.p, nullptr, tpart.p, s);
    dot_and_sum(y.p, alpha.p, logdiag.p, n, scal.p, s);
    GPRX_HIP_CHECK(hipEventRecord(ev[3], s));
    if (want_grad) {
      gemm_call<double>(true, false, X64.p, npad, X64.p, npad, A64.p, npad, (int)npad, (int)npad,
                        (int)npad, 1.0, 0.0, kKloRow, 1, kOrderRowFast, s);
      GPRX_HIP_CHECK(hipEventRecord(ev[4], s));
      contract_exact<double>(dp, hp.ns, prog.p, pts(), A64.p, npad, alpha.p, partials.p, s);
      reduce_partials(partials.p, lower_tiles(npad), accs.p, s);
      noise_trace<double>(alpha.p, A64.p, npad, n, scal.p + 2, s);
    } else {
      GPRX_HIP_CHECK(hipEventRecord(ev[4], s));
    }
    GPRX_HIP_CHECK(hipEventRecord(ev[5], s));
    cur_form = GPRX_FP64;
    return readback(3, want_grad);
  }

  // MIXED form.  Returns info bits, or -1 when refinement did not converge.
  int run_mixed(bool want_grad) {
    alloc_mixed();
    hipStream_t s = stream;
    GPRX_HIP_CHECK(hipMemsetAsync(info.p, 0, sizeof(int), s));
    fill_lower<float>(dp, prog.p, pts(), A32.p, npad, K64.p, info.p, s);
    GPRX_HIP_CHECK(hipEventRecord(ev[1], s));
    factor_inverse<float>(A32.p, X32.p, T32.p, npad, ldt, nt, logdiag.p, info.p, fstreams);
    GPRX_HIP_CHECK(hipEventRecord(ev[2], s));
    tile_mv_lower<float>(X32.p, npad, nt, kTmvRow, y.p, tv.p, nullptr, tpart.p, s);
    tile_mv_lower<float>(X32.p, npad, nt, kTmvCol, tv.p, alpha.p, nullptr, tpart.p, s);
    int bits = readback(0, false);
    if (bits) return bits;
    // Iterative refinement (src/precision/refine.rs:56-98): residual in f64
    // against the f64 lower triangle, corrections through the f32 inverse.
    const double tol = 10.0 * (double)n * 2.220446049250313e-16;
    double prev = -1.0;
    int streak = 0;
    bool ok = false;
    refine_steps = 0;
    for (int it = 0; it < 10; ++it) {
      // A alpha from the f64 lower triangle; on the first pass also the
      // absolute row sums, whose maximum is ||A||_inf
      tile_mv_lower<double>(K64.p, npad, nt, it == 0 ? kTmvSymAbs : kTmvSym, alpha.p, p1.p, q1.p, tpart.p, s);
      residual_norms(y.p, p1.p, nullptr, alpha.p, it == 0 ? q1.p : nullptr, nullptr, r.p, n, scal.p + 4, s);
      readback(8, false);
      const double rinf = hscal.p[4], winf = hscal.p[5], yinf = hscal.p[6], ainf = hscal.p[7];
      const double denom = ainf * winf + yinf;
      if (denom > 0.0 && rinf / denom < tol) {
        ok = true;
        break;
      }
      if (prev >= 0.0) {
        const double ratio = prev > 0.0 ? rinf / prev : 0.0;
        if (ratio > 0.9) {
          if (++streak >= 2) break;
        } else {
          streak = 0;
        }
      }
      prev = rinf;
      tile_mv_lower<float>(X32.p, npad, nt, kTmvRow, r.p, tv.p, nullptr, tpart.p, s);
      tile_mv_lower<float>(X32.p, npad, nt, kTmvCol, tv.p, dlt.p, nullptr, tpart.p, s);
      axpy(alpha.p, dlt.p, n, s);
      ++refine_steps;
    }
    if (!ok) return -1;
    dot_and_sum(y.p, alpha.p, logdiag.p, n, scal.p, s);
    GPRX_HIP_CHECK(hipEventRecord(ev[3], s));
    if (want_grad) {
      gemm_call<float>(true, false, X32.p, npad, X32.p, npad, A32.p, npad, (int)npad, (int)npad,
                       (int)npad, 1.0, 0.0, kKloRow, 1, kOrderRowFast, s);
      GPRX_HIP_CHECK(hipEventRecord(ev[4], s));
      contract_exact<float>(dp, hp.ns, prog.p, pts(), A32.p, npad, alpha.p, partials.p, s);
      reduce_partials(partials.p, lower_tiles(npad), accs.p, s);
      noise_trace<float>(alpha.p, A32.p, npad, n, scal.p + 2, s);
    } else {
      GPRX_HIP_CHECK(hipEventRecord(ev[4], s));
    }
    GPRX_HIP_CHECK(hipEventRecord(ev[5], s));
    cur_form = GPRX_MIXED;
    return readback(3, want_grad);
  }

  // One attempt at jitter j in the mixed form (FP64 when refinement stalls)
  // or in FP64; fills nlml/grad on success.  Returns info bits.
  int attempt(const double* th, double j, bool fp64, bool want_grad, double* nlml, double* grad) {
    upload_program(th, j);
    GPRX_HIP_CHECK(hipEventRecord(ev[0], stream));
    int bits;
    if (!fp64) {
      bits = run_mixed(want_grad);
      if (bits == -1) {
        fallback = GPRX_FALLBACK_REFINEMENT;
        bits = run_fp64(want_grad);
      }
    } else {
      bits = run_fp64(want_grad);
    }
    if (bits == 0) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[0], ev[5]);
      last_eval_ms = ms;
      for (int k = 0; k < 5; ++k) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, ev[k], ev[k + 1]);
        phase_ms[k] = t;
      }
      last_dense_ms = phase_ms[1];
    }
    if (bits) return bits;
    const double* sc = hscal.p;
    *nlml = 0.5 * sc[0] + sc[1] + 0.5 * (double)n * kLog2Pi;
    if (!std::isfinite(*nlml)) return 2;
    if (want_grad && grad) {
      acc_to_grad(hp, th, hacc.p, grad);
      for (int p = 0; p < hp.n_params; ++p) grad[p] *= -0.5;
      grad[hp.n_params] = -0.5 * std::exp(th[hp.n_params]) * sc[2];
    }
    return 0;
  }

  // The jitter loop of src/linalg/chol.rs:20-43 around one evaluation.  In
  // the mixed form, a failed f32 factor whose next jitter would pass
  // mixed_jitter_bound (or the policy's end) is formed again in FP64 at the
  // same jitter, and the loop goes on in FP64.
  int evaluate(const double* th, bool want_grad, double* nlml, double* grad, gprx_error* err) {
    z_valid = false;  // X changes
    double j = 0.0, next = jit.initial;
    int remaining = jit.max_retries;
    bool fp64 = precision != GPRX_MIXED;
    fallback = GPRX_FALLBACK_NONE;
    for (;;) {
      const int bits = attempt(th, j, fp64, want_grad, nlml, grad);
      if (bits == 0) {
        last_jitter = j;
        factored = true;
        return GPRX_OK;
      }
      factored = false;
      if (bits & 2) {
        set_error(err, GPRX_NON_FINITE_KERNEL_VALUE, "kernel matrix holds a non-finite value");
        return GPRX_NON_FINITE_KERNEL_VALUE;
      }
      const bool exhausted = remaining == 0 || !(next > 0.0) || next > jit.max_jitter;
      if (!fp64 && (exhausted || next > mixed_jitter_bound)) {
        fp64 = true;
        fallback = GPRX_FALLBACK_JITTER;
        continue;
      }
      if (exhausted) {
        set_error(err, GPRX_CHOLESKY_FAILED, "Cholesky factor failed", GPRX_STAGE_FIT, j, (size_t)n);
        return GPRX_CHOLESKY_FAILED;
      }
      --remaining;
      j = next;
      next *= jit.multiplier;
      if (!std::isfinite(next)) remaining = 0;
    }
  }

  // NLML, gradient and Hessian at th in FP64 (ExactFit::fill_hessian_from_factor,
  // src/gpr/exact_fit.rs:484-596; the pieces are described in hessian.hip).
  // H is p x p, row-major.  Works on buffers of its own: the model's factor,
  // alpha and inverse stay those of th afterwards.
  int hessian(const double* th, double* nlml, double* grad, double* H, gprx_error* err) {
    const int saved = precision;
    precision = GPRX_FP64;
    int rc = evaluate(th, true, nlml, grad, err);
    precision = saved;
    if (rc != GPRX_OK) return rc;
    hipStream_t s = stream;
    const int nk = hp.n_params, np = nk + 1;
    const double sigma2 = std::exp(th[nk]);
    // second order: kernel pairs by contraction, the noise pair from the
    // noise trace (d2A/dlog sigma^2 2 = sigma^2 I), no kernel-noise cross term
    std::vector<double> h2((size_t)np * np, 0.0);
    const int tiles = lower_tiles(npad);
    DevBuf<double> part2;
    HostBuf<double> hpart;
    part2.alloc((size_t)tiles);
    hpart.alloc((size_t)std::max<long>(tiles, npad));
    const HessSlots hs = hess_slots(hp);
    for (int i = 0; i < nk; ++i)
      for (int j = i; j < nk; ++j) {
        contract2(dp, prog.p, hs, pts(), A64.p, npad, alpha.p, i, j, part2.p, s);
        GPRX_HIP_CHECK(hipMemcpyAsync(hpart.p, part2.p, sizeof(double) * tiles,
                                      hipMemcpyDeviceToHost, s));
        GPRX_HIP_CHECK(hipStreamSynchronize(s));
        double acc = 0.0;
        for (int t = 0; t < tiles; ++t) acc += hpart.p[t];
        h2[(size_t)i * np + j] = h2[(size_t)j * np + i] = -0.5 * acc;
      }
    h2[(size_t)nk * np + nk] = -0.5 * sigma2 * hscal.p[2];
    // first order: S_i = X A_i X^T (lower), v_i = X (A_i alpha)
    const size_t nn = (size_t)npad * npad;
    DevBuf<double> D, Y, S, u, v, gp;
    D.alloc(nn);
    Y.alloc(nn);
    S.alloc(nn * np);
    u.alloc((size_t)npad);
    v.alloc((size_t)npad * np);
    gp.alloc((size_t)npad * ((npad + 255) / 256));
    GPRX_HIP_CHECK(hipMemsetAsync(D.p, 0, sizeof(double) * nn, s));
    for (int i = 0; i < np; ++i) {
      if (i < nk) {
        dfill(dp, prog.p, hs, pts(), i, D.p, npad, s);
      } else {
        GPRX_HIP_CHECK(hipMemsetAsync(D.p, 0, sizeof(double) * nn, s));
        set_diag(D.p, npad, n, sigma2, s);
      }
      gemv_n<double>(D.p, npad, (int)npad, (int)npad, alpha.p, u.p, gp.p, 0, 0, 0, s);
      gemv_n<double>(X64.p, npad, (int)npad, (int)npad, u.p, v.p + (size_t)i * npad, gp.p, 1, 0, 0,
                     s);
      gemm_call<double>(false, false, X64.p, npad, D.p, npad, Y.p, npad, (int)npad, (int)npad,
                        (int)npad, 1.0, 0.0, kKhiRow, 0, kOrderRowFast, s);
      gemm_call<double>(false, true, Y.p, npad, X64.p, npad, S.p + (size_t)i * nn, npad, (int)npad,
                        (int)npad, (int)npad, 1.0, 0.0, kKhiCol, 1, kOrderRowFast, s);
    }
    HostBuf<double> hv;
    hv.alloc((size_t)npad * np);
    GPRX_HIP_CHECK(hipMemcpyAsync(hv.p, v.p, sizeof(double) * npad * np, hipMemcpyDeviceToHost, s));
    DevBuf<double> pd;
    pd.alloc((size_t)npad);
    for (int i = 0; i < np; ++i)
      for (int j = 0; j <= i; ++j) {
        dot_lower(S.p + (size_t)i * nn, S.p + (size_t)j * nn, npad, n, pd.p, s);
        GPRX_HIP_CHECK(hipMemcpyAsync(hpart.p, pd.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        GPRX_HIP_CHECK(hipStreamSynchronize(s));
        double tr = 0.0, vv = 0.0;
        for (int c = 0; c < n; ++c) tr += hpart.p[c];
        for (int r = 0; r < n; ++r) vv += hv.p[(size_t)i * npad + r] * hv.p[(size_t)j * npad + r];
        const double h = h2[(size_t)i * np + j] - 0.5 * tr + vv;
        H[(size_t)i * np + j] = H[(size_t)j * np + i] = h;
      }
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    return GPRX_OK;
  }
};

int validate_theta(const HostProgram& hp, const double* th, size_t p, gprx_error* err) {
  if (p != (size_t)hp.n_params + 1) {
    set_error(err, GPRX_LENGTH_MISMATCH, "parameter count differs from the program's");
    return GPRX_LENGTH_MISMATCH;
  }
  for (size_t i = 0; i < p; ++i)
    if (!std::isfinite(th[i])) {
      set_error(err, i + 1 == p ? GPRX_INVALID_NOISE_VARIANCE : GPRX_INVALID_HYPERPARAMETER,
                "non-finite log parameter");
      return i + 1 == p ? GPRX_INVALID_NOISE_VARIANCE : GPRX_INVALID_HYPERPARAMETER;
    }
  const double noise = std::exp(th[p - 1]);
  if (!(noise > 0.0) || !std::isfinite(noise)) {
    set_error(err, GPRX_INVALID_NOISE_VARIANCE, "noise variance must be positive and finite");
    return GPRX_INVALID_NOISE_VARIANCE;
  }
  return GPRX_OK;
}

template <typename F>
int guarded(gprx_error* err, F&& f) {
  try {
    return f();
  } catch (const DeviceError& e) {
    set_error(err, GPRX_DEVICE_ERROR, e.msg);
    return GPRX_DEVICE_ERROR;
  } catch (const std::bad_alloc&) {
    set_error(err, GPRX_DEVICE_ERROR, "host allocation failed");
    return GPRX_DEVICE_ERROR;
  } catch (...) {
    set_error(err, GPRX_DEVICE_ERROR, "unexpected failure");
    return GPRX_DEVICE_ERROR;
  }
}

// Uploads column-major points (n x d) padded to (rows_pad x dp).
void upload_points(DevBuf<double>& dst, const double* src, int n, int d, long rows_pad, int dp,
                   hipStream_t s) {
  std::vector<double> h((size_t)rows_pad * dp, 0.0);
  for (int k = 0; k < d; ++k)
    for (int i = 0; i < n; ++i) h[(size_t)k * rows_pad + i] = src[(size_t)k * n + i];
  dst.alloc(h.size());
  GPRX_HIP_CHECK(hipMemcpyAsync(dst.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s));
  GPRX_HIP_CHECK(hipStreamSynchronize(s));
}

// The same from point-major host data (point i at src[i d .. i d + d)).
void upload_point_rows(DevBuf<double>& dst, const double* src, int n, int d, long rows_pad, int dp,
                       hipStream_t s) {
  std::vector<double> h((size_t)rows_pad * dp, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < d; ++k) h[(size_t)k * rows_pad + i] = src[(size_t)i * d + k];
  dst.alloc(h.size());
  GPRX_HIP_CHECK(hipMemcpyAsync(dst.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s));
  GPRX_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

namespace gprx {
int exact_validate_theta(const HostProgram& hp, const double* th, size_t p, gprx_error* err) {
  return validate_theta(hp, th, p, err);
}
void exact_upload_points(DevBuf<double>& dst, const double* src, int n, int d, long rows_pad, int dp,
                         hipStream_t s) {
  upload_points(dst, src, n, d, rows_pad, dp, s);
}
}  // namespace gprx

// ---------------------------------------------------------------------------
// Prediction without allocation (gprx_hip_exact_predict_noalloc).  Its work
// space is one file-scope device array, `scratch`, of kNaScratchDoubles
// doubles (32 MiB; each GPU the library runs on holds one until the process
// ends).  Queries run in chunks of kNaCols columns, so the array bounds the
// model's padded point count (kNaRows), not the number of queries:
//   [0, 4096)               the chunk's query points, kNaCols x 32 dimensions
//   [4096, +rows*cols)       K(x, x*) of the chunk, npad x kNaCols (f64)
//   [.., +rows*cols)         V = L^-1 K(x, x*) (f64), or its f32 K and V
//   [.., +5*kNaCols)         k(x*, x*), column sums of squares, mean, and the
//                            chunk's mean and variance staged for the host
// Host code never uses the array's name as an address: hipGetSymbolAddress
// gives each device's copy once (kept in g_na_ptr), and every kernel
// argument and copy uses that pointer.
constexpr long kNaRows = 16384;
constexpr long kNaCols = kTile;
constexpr long kNaDims = 32;
constexpr long kNaOffX = 0;
constexpr long kNaOffK = kNaCols * kNaDims;
constexpr long kNaOffV = kNaOffK + kNaRows * kNaCols;
constexpr long kNaOffVec = kNaOffV + kNaRows * kNaCols;
constexpr long kNaScratchDoubles = kNaOffVec + 5 * kNaCols;
static_assert((kNaOffV * sizeof(double)) % 256 == 0, "f32 views of V start 256-byte aligned");

static __device__ __align__(256) double scratch[kNaScratchDoubles];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_na_init_mu;
double* g_na_ptr[kNaMaxDevices] = {};
// One call at a time per device: calls share that device's array.
std::mutex g_na_call_mu[kNaMaxDevices];

// The current device's copy of `scratch` (looked up once per device).
int na_scratch(int* dev_out, double** out, gprx_error* err) {
  int dev = 0;
  GPRX_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDevices) {
    set_error(err, GPRX_UNSUPPORTED, "device index beyond the no-allocation table");
    return GPRX_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_na_init_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(err, GPRX_DEVICE_ERROR, "hipGetSymbolAddress of the prediction scratch failed");
      return GPRX_DEVICE_ERROR;
    }
    g_na_ptr[dev] = static_cast<double*>(p);
  }
  *dev_out = dev;
  *out = g_na_ptr[dev];
  return GPRX_OK;
}

// The chunk's outputs: mean, and the variance clamped at 0 (src/gpr/shared.rs:798-806)
// with the noise added unless latent.
__global__ void k_na_finish(const double* __restrict__ mu, const double* __restrict__ kss,
                            const double* __restrict__ ss, int qc, int want_var, int latent,
                            double noise, double* __restrict__ out_mean, double* __restrict__ out_var) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= qc) return;
  out_mean[j] = mu[j];
  if (want_var) {
    double lat = kss[j] - ss[j];
    if (lat < 0.0) lat = 0.0;
    out_var[j] = latent ? lat : lat + noise;
  }
}

}  // namespace

extern "C" {

const char* gprx_hip_version(void) { return "gprx-hip 0.1.0 (gfx950)"; }

int gprx_hip_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

// Buffers sized by npad, beyond the points and targets (set-up, and again
// when online insertion outgrows the padding).
void size_buffers(ExactModel* m) {
  hipStream_t s = m->stream;
  m->prog.alloc(1);
  m->hprog.alloc(1);
  const long np = m->npad;
  m->logdiag.alloc(np);
  m->alpha.alloc(np);
  m->tv.alloc(np);
  m->r.alloc(np);
  m->dlt.alloc(np);
  m->p1.alloc(np);
  m->q1.alloc(np);
  m->partials.alloc((size_t)lower_tiles(np) * kMaxAcc);
  m->tpart.alloc((size_t)lower_tiles(np) * 4 * kTile);
  m->accs.alloc(kMaxAcc);
  m->scal.alloc(16);
  m->info.alloc(1);
  m->hscal.alloc(16);
  m->hacc.alloc(kMaxAcc);
  m->hinfo.alloc(1);
  GPRX_HIP_CHECK(hipMemsetAsync(m->r.p, 0, sizeof(double) * np, s));
  GPRX_HIP_CHECK(hipMemsetAsync(m->alpha.p, 0, sizeof(double) * np, s));
  GPRX_HIP_CHECK(hipMemsetAsync(m->logdiag.p, 0, sizeof(double) * np, s));
  GPRX_HIP_CHECK(hipMemsetAsync(m->scal.p, 0, sizeof(double) * 16, s));
  if (m->precision == GPRX_MIXED)
    m->alloc_mixed();
  else
    m->alloc_fp64();
}

// Scalars of one insertion from sc[8] = l^T l, sc[10] = l^T z and
// sc[12] = k(x_new, x_new): sc[13] = 1 / d, sc[14] = z_n, sc[15] = 1 when
// the pivot d^2 = kappa - l^T l is positive (0 otherwise, and nothing below
// changes the model); z[n] = z_n and logdiag[n] = log d.
// out = the lower triangle of in^T, zeros above the diagonal (both ld x ld
// column-major, ld a multiple of 64): one 64 x 64 tile per block, read and
// written along columns through a padded LDS tile.
__global__ __launch_bounds__(256) void k_transpose_lower(const double* __restrict__ in,
                                                         double* __restrict__ out, long ld) {
  __shared__ double t[64][65];
  const long i0 = (long)blockIdx.x * 64, j0 = (long)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int k = ty; k < 64; k += 4) t[k][tx] = in[(i0 + tx) + (j0 + k) * ld];
  __syncthreads();
  // out(j, i) = in(i, j) = t[j - j0][i - i0]
  for (int k = ty; k < 64; k += 4) {
    const long i = i0 + k, j = j0 + tx;
    out[j + i * ld] = j >= i ? t[tx][k] : 0.0;
  }
}

// The inserted point by value (kernel argument): no host staging.
struct NewPoint {
  double x[kMaxDim];
  double y;
};

// xnew (kTile x dp, the one-point set) = the point in row 0, zeros below;
// row n of the model's x and y = the point and its target.
__global__ void k_insert_point(NewPoint pt, double* __restrict__ xnew, double* __restrict__ x,
                               long ldx, double* __restrict__ y, int n, int d) {
  const int q = blockIdx.x, i = threadIdx.x;
  const double v = q < d ? pt.x[q] : 0.0;
  xnew[(long)q * kTile + i] = i == 0 ? v : 0.0;
  if (i == 0 && q < d) x[(long)q * ldx + n] = v;
  if (i == 0 && q == 0) y[n] = pt.y;
}

__global__ void k_insert_scalars(double* sc, double noise, double y_new, double* z, double* logdiag,
                                 int n) {
  const double piv2 = sc[12] + noise - sc[8];
  const bool ok = piv2 > 0.0 && isfinite(piv2);
  const double d = ok ? sqrt(piv2) : 1.0;
  sc[13] = 1.0 / d;
  sc[14] = ok ? (y_new - sc[10]) / d : 0.0;
  sc[15] = ok ? 1.0 : 0.0;
  if (ok) {
    z[n] = sc[14];
    logdiag[n] = log(d);
  }
}

// Row n of X (the new point's): X[n][j] = -w_j / d for j < n, X[n][n] = 1 / d
// (1 / d = sc[13]; only when sc[15] says the pivot held).
__global__ void k_insert_row(double* X, long ld, int n, const double* __restrict__ w,
                             const double* __restrict__ sc) {
  if (sc[15] == 0.0) return;
  const double inv_d = sc[13];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) X[n + (long)j * ld] = -w[j] * inv_d;
  if (j == n) X[n + (long)n * ld] = inv_d;
}

// alpha after a one-row extension of X: alpha_j -= c w_j (j < n), alpha_n = c,
// with c = z_n / d and w = X^T l (the new row of X is -w / d, 1 / d).
__global__ void k_alpha_extend(double* alpha, const double* __restrict__ w, int n,
                               const double* __restrict__ sc) {
  if (sc[15] == 0.0) return;
  const double c = sc[14] * sc[13];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) alpha[j] -= c * w[j];
  if (j == n) alpha[n] = c;
}

// Deletion of point p from the factor, on X = L^-1 alone (OnlineGpr::delete,
// src/gpr/online.rs:564-620, which downdates its LDL^T, src/linalg/ldlt.rs:113).
// With L = [L11 0 0; r^T d 0; L31 l L33], removing row and column p leaves
// L31 and L11 and turns L33 into L33' = L33 G, G G^T = I + v v^T with
// v = L33^-1 l = -X[p+1:, p] / X[p][p]: a rank-one update, always defined.
// In terms of X, rows above p keep X11 and the trailing rows become
//   [X31', X33'] = G^-1 ([X31, X33] + v [X[p, :p], 0]),
// and G^-1 applies as a prefix sum per column: with beta_k = 1 + sum_{i<=k} v_i^2,
//   y_k = (m_k - v_k S_k / beta_{k-1}) / G_kk,  S_k = sum_{i<k} v_i m_i,
//   G_kk = sqrt(beta_k / beta_{k-1}).
// One workgroup per new column jn (old column jn, or jn + 1 from p on) reads
// X and writes the shifted matrix into Y; the last new row and column are
// padding (identity).  v, 1 / beta_{k-1} and 1 / G_kk arrive from the host.
__global__ __launch_bounds__(256) void k_delete_rows(const double* __restrict__ X,
                                                     double* __restrict__ Y, long ld, int p,
                                                     const double* __restrict__ v,
                                                     const double* __restrict__ ibp,
                                                     const double* __restrict__ ig) {
  const int jn = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double* yc = Y + (long)jn * ld;
  if (jn == ld - 1) {  // new padding column: e_{ld-1}
    for (long i = tid; i < ld; i += 256) yc[i] = (i == ld - 1) ? 1.0 : 0.0;
    return;
  }
  const long j = jn < p ? jn : jn + 1;
  const double* xc = X + j * ld;
  // rows above p: old column j < p copied from its diagonal down, zeros
  // above it (X's upper triangle is zero; the export's X K + s X reads it)
  for (int i = tid; i < p; i += 256) yc[i] = (j < p && i >= j) ? xc[i] : 0.0;
  if (tid == 0) yc[ld - 1] = 0.0;  // new padding row
  const int mt = (int)(ld - 1 - p);  // trailing rows: new p .. ld-2 from old p+1 .. ld-1
  if (mt <= 0) return;
  const double xp = (j < p) ? X[p + j * ld] : 0.0;
  // above old row j of an old column j > p the column is zero, and so is y
  const int k0 = (j > p) ? (int)(j - (p + 1)) : 0;
  for (int k = tid; k < k0; k += 256) yc[p + k] = 0.0;
  __shared__ double wsum[4];
  constexpr int E = 4;  // consecutive rows per thread per chunk
  double carry = 0.0;
  for (int c0 = k0; c0 < mt; c0 += 256 * E) {
    double mv[E], pr[E];
    double tsum = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = c0 + tid * E + e;
      double val = 0.0, prod = 0.0;
      if (k < mt) {
        val = xc[p + 1 + k] + v[k] * xp;
        prod = v[k] * val;
      }
      mv[e] = val;
      pr[e] = prod;
      tsum += prod;
    }
    // exclusive prefix of the threads' sums over the workgroup (fixed order)
    double incl = tsum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    double before = 0.0;
    for (int w = 0; w < wave; ++w) before += wsum[w];
    const double chunk = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    double run = carry + before + (incl - tsum);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = c0 + tid * E + e;
      if (k < mt) yc[p + k] = (mv[e] - v[k] * run * ibp[k]) * ig[k];
      run += pr[e];
    }
    carry += chunk;
  }
}

// log L'_ii: rows above p keep theirs; trailing rows gain log G_kk; the last
// row is padding (log 1).
__global__ void k_delete_logdiag(const double* __restrict__ ld_old, const double* __restrict__ lg,
                                 double* __restrict__ out, int p, long np) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  if (i < p)
    out[i] = ld_old[i];
  else if (i < np - 1)
    out[i] = ld_old[i + 1] + lg[i - p];
  else
    out[i] = 0.0;
}

// Coefficients of the downdate of point p (k_delete_rows): v_k =
// -X[p+1+k][p] / X[p][p], beta_k = 1 + sum_{i<=k} v_i^2 by a fixed-order
// scan, 1 / beta_{k-1}, 1 / G_kk and log G_kk = 1/2 log1p(v_k^2 / beta_{k-1})
// for k < mt.  One workgroup.
constexpr int kCoeffThreads = 1024;
__global__ __launch_bounds__(kCoeffThreads) void k_delete_coeffs(const double* __restrict__ X,
                                                                 long ld, int p, int mt,
                                                                 double* __restrict__ v,
                                                                 double* __restrict__ ibp,
                                                                 double* __restrict__ ig,
                                                                 double* __restrict__ lg) {
  constexpr int W = kCoeffThreads / 64, E = 16;
  __shared__ double wsum[W];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double* col = X + (long)p * ld + p;  // col[0] = X[p][p]
  const double xpp = col[0];
  double carry = 0.0;
  for (int c0 = 0; c0 < mt; c0 += kCoeffThreads * E) {
    double sq[E];
    double tsum = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = c0 + tid * E + e;
      double vk = 0.0;
      if (k < mt) {
        vk = -col[1 + k] / xpp;
        v[k] = vk;
      }
      sq[e] = vk * vk;
      tsum += sq[e];
    }
    double incl = tsum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    double before = 0.0, chunk = 0.0;
    for (int w = 0; w < W; ++w) {
      if (w < wave) before += wsum[w];
      chunk += wsum[w];
    }
    __syncthreads();
    double run = carry + before + (incl - tsum);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = c0 + tid * E + e;
      if (k < mt) {
        const double bp = 1.0 + run, b = bp + sq[e];
        ibp[k] = 1.0 / bp;
        ig[k] = 1.0 / sqrt(b / bp);
        lg[k] = 0.5 * log1p(sq[e] / bp);
      }
      run += sq[e];
    }
    carry += chunk;
  }
}

// Rows p+1 .. n-1 of the points (dp columns of ld rows) and of y, out to
// tmp; k_shift_rows_in writes them back one row up and clears row n-1, so
// no thread reads a row another has moved.
__global__ void k_shift_rows_out(const double* __restrict__ x, long ld, int dp,
                                 const double* __restrict__ y, int n, int p, double* __restrict__ tmp) {
  const long cnt = n - 1 - p, total = cnt * (dp + 1);
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long q = idx / cnt, k = idx % cnt;
    tmp[idx] = q < dp ? x[q * ld + p + 1 + k] : y[p + 1 + k];
  }
}
__global__ void k_shift_rows_in(double* __restrict__ x, long ld, int dp, double* __restrict__ y, int n,
                                int p, const double* __restrict__ tmp) {
  const long cnt = n - 1 - p, total = (cnt + 1) * (dp + 1);
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long q = idx / (cnt + 1), k = idx % (cnt + 1);
    const double val = k < cnt ? tmp[q * cnt + k] : 0.0;
    if (q < dp)
      x[q * ld + p + k] = val;
    else
      y[p + k] = val;
  }
}

// Re-pads the model for n_new points: every npad-sized buffer again, the
// points and targets from the host copies, no factor.
void regrow(ExactModel* m, long npad_new) {
  hipStream_t s = m->stream;
  GPRX_HIP_CHECK(hipStreamSynchronize(s));
  m->npad = npad_new;
  m->nt = (int)(npad_new / kTile);
  m->ldt = (long)((m->nt + 1) / 2) * kTile;
  for (DevBuf<double>* b : {&m->A64, &m->X64, &m->T64, &m->K64}) b->release();
  for (DevBuf<float>* b : {&m->A32, &m->X32, &m->T32}) b->release();
  size_buffers(m);
  upload_point_rows(m->x, m->hx.data(), m->n, m->d, m->npad, m->dp, s);
  std::vector<double> yh(m->npad, 0.0);
  std::memcpy(yh.data(), m->hy.data(), sizeof(double) * m->n);
  m->y.alloc(m->npad);
  GPRX_HIP_CHECK(hipMemcpyAsync(m->y.p, yh.data(), sizeof(double) * m->npad, hipMemcpyHostToDevice, s));
  GPRX_HIP_CHECK(hipStreamSynchronize(s));
  m->factored = false;
}

// OnlineGpr::insert (src/gpr/online.rs) on the FP64 factor, O(n^2): with
// k = K(X, x_new), kappa = k(x_new, x_new) + noise and l = L^-1 k = X k,
// the new row of L is [l^T, d], d = sqrt(kappa - l^T l); the new row of
// X = L^-1 is [-l^T X / d, 1 / d]; z = X y gains (y_new - l^T z) / d, and
// alpha = X^T z = alpha_old - (z_n / d) X^T l (plus alpha_n = z_n / d),
// NLML = 1/2 z^T z + sum log L_ii + n/2 log 2 pi.  z is kept from one
// insertion to the next (recomputed after a full evaluation), so an
// insertion reads X's lower triangle twice (X k, X^T l), one workgroup per
// 128-tile.  When the padding
// is full the model is re-padded by 128 rows and factored again.
int exact_insert(ExactModel* m, const double* xn, double yn, double* nlml, gprx_error* err) {
  if (m->precision != GPRX_FP64) {
    set_error(err, GPRX_UNSUPPORTED, "online insertion runs on FP64 models");
    return GPRX_UNSUPPORTED;
  }
  if (!all_finite(xn, (size_t)m->d) || !std::isfinite(yn)) {
    set_error(err, GPRX_NON_FINITE_INPUT, "non-finite point");
    return GPRX_NON_FINITE_INPUT;
  }
  hipStream_t s = m->stream;
  const int n = m->n, d = m->d;
  const double* th = m->theta.data();
  if (n + 1 > m->npad) {
    m->hx.insert(m->hx.end(), xn, xn + d);
    m->hy.push_back(yn);
    m->n = n + 1;
    regrow(m, m->npad + kTile);
    double v = 0.0;
    int rc = m->evaluate(th, false, &v, nullptr, err);
    if (rc != GPRX_OK) {  // back to the model before the insert
      m->hx.resize((size_t)n * d);
      m->hy.pop_back();
      m->n = n;
      regrow(m, m->npad - kTile);
      gprx_error e2{};
      double v2 = 0.0;
      (void)m->evaluate(th, false, &v2, nullptr, &e2);
      return rc;
    }
    if (nlml) *nlml = v;
    return GPRX_OK;
  }
  if (!m->factored || m->cur_form != GPRX_FP64) {
    double v = 0.0;
    int rc = m->evaluate(th, false, &v, nullptr, err);
    if (rc != GPRX_OK) return rc;
  }
  const long np = m->npad;
  if (m->zv.n != (size_t)np) m->z_valid = false;
  m->zv.alloc(np);
  m->kcol.alloc((size_t)np * kTile);
  m->lv.alloc(np);
  m->wv.alloc(np);
  m->xnew.alloc((size_t)kTile * m->dp);
  // the new point as a one-row point set (and its row of x and y, cleared
  // again below if the pivot fails), and z = X y of the current factor
  NewPoint pt{};
  for (int q = 0; q < d; ++q) pt.x[q] = xn[q];
  pt.y = yn;
  hipLaunchKernelGGL(k_insert_point, dim3(m->dp), dim3(kTile), 0, s, pt, m->xnew.p, m->x.p, np,
                     m->y.p, n, d);
  GPRX_HIP_CHECK(hipGetLastError());
  const DevPoints pn{m->xnew.p, kTile, 1};
  // Products with X run over its padded lower tiles: the padding block of X
  // is the identity and the padded entries of y and of k are zero, so the
  // first n entries are those of the n x n products.
  if (!m->z_valid)
    tile_mv_lower<double>(m->X64.p, np, m->nt, kTmvRow, m->y.p, m->zv.p, nullptr, m->tpart.p, s);
  fill_cross<double>(m->dp, m->prog.p, m->pts(), pn, m->kcol.p, np, s);
  fill_diag(m->dp, m->prog.p, pn, m->scal.p + 12, s);
  tile_mv_lower<double>(m->X64.p, np, m->nt, kTmvRow, m->kcol.p, m->lv.p, nullptr, m->tpart.p, s);
  dot_and_sum(m->lv.p, m->lv.p, m->lv.p, n, m->scal.p + 8, s);
  dot_and_sum(m->lv.p, m->zv.p, m->lv.p, n, m->scal.p + 10, s);
  // the pivot, z_n and log d on the device; every update below is skipped
  // there when the pivot fails, so the host reads back once
  const double noise = std::exp(th[m->hp.n_params]) + m->last_jitter;
  hipLaunchKernelGGL(k_insert_scalars, dim3(1), dim3(1), 0, s, m->scal.p, noise, yn, m->zv.p,
                     m->logdiag.p, n);
  GPRX_HIP_CHECK(hipGetLastError());
  // X's new row, then alpha = X^T z: the old alpha (= X^T z of the old
  // rows) plus the new row's term, from w = X^T l; the NLML sums
  tile_mv_lower<double>(m->X64.p, np, m->nt, kTmvCol, m->lv.p, m->wv.p, nullptr, m->tpart.p, s);
  hipLaunchKernelGGL(k_insert_row, dim3((n + 1 + 255) / 256), dim3(256), 0, s, m->X64.p, np, n,
                     m->wv.p, m->scal.p);
  GPRX_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_alpha_extend, dim3((n + 1 + 255) / 256), dim3(256), 0, s, m->alpha.p,
                     m->wv.p, n, m->scal.p);
  GPRX_HIP_CHECK(hipGetLastError());
  dot_and_sum(m->zv.p, m->zv.p, m->logdiag.p, n + 1, m->scal.p, s);
  GPRX_HIP_CHECK(hipMemcpyAsync(m->hscal.p, m->scal.p, sizeof(double) * 16, hipMemcpyDeviceToHost, s));
  GPRX_HIP_CHECK(hipStreamSynchronize(s));
  if (m->hscal.p[15] == 0.0) {
    // the point's row of x and its target lie in the padding: clear them
    const std::vector<double> zero((size_t)d, 0.0);
    GPRX_HIP_CHECK(hipMemcpy2DAsync(m->x.p + n, sizeof(double) * np, zero.data(), sizeof(double),
                                    sizeof(double), d, hipMemcpyHostToDevice, s));
    const double z0 = 0.0;
    GPRX_HIP_CHECK(hipMemcpyAsync(m->y.p + n, &z0, sizeof(double), hipMemcpyHostToDevice, s));
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    set_error(err, GPRX_CHOLESKY_FAILED, "Cholesky factor failed on insertion", GPRX_STAGE_FIT,
              m->last_jitter, (size_t)(n + 1));
    return GPRX_CHOLESKY_FAILED;
  }
  m->n = n + 1;
  m->hx.insert(m->hx.end(), xn, xn + d);
  m->hy.push_back(yn);
  m->z_valid = true;
  if (nlml) *nlml = 0.5 * m->hscal.p[0] + m->hscal.p[1] + 0.5 * (double)(n + 1) * kLog2Pi;
  return GPRX_OK;
}

int gprx_hip_exact_create(const double* x, size_t n, size_t d, const double* y,
                          const gprx_kernel_program* kernel, const double* log_theta, size_t p,
                          int32_t precision, const gprx_jitter_policy* jitter,
                          gprx_hip_model** out, gprx_error* err) {
  if (out) *out = nullptr;
  return guarded(err, [&]() -> int {
    if (!x || !y || !out || n == 0 || d == 0) {
      set_error(err, GPRX_EMPTY_INPUT, "empty training data");
      return GPRX_EMPTY_INPUT;
    }
    if (d > (size_t)kMaxDim) {
      set_error(err, GPRX_UNSUPPORTED, "input dimension above 32");
      return GPRX_UNSUPPORTED;
    }
    if (n > (size_t)1 << 20) {
      set_error(err, GPRX_UNSUPPORTED, "exact GPR above 2^20 points");
      return GPRX_UNSUPPORTED;
    }
    if (!all_finite(x, n * d) || !all_finite(y, n)) {
      set_error(err, GPRX_NON_FINITE_INPUT, "non-finite training data");
      return GPRX_NON_FINITE_INPUT;
    }
    if (precision != GPRX_FP64 && precision != GPRX_MIXED) {
      set_error(err, GPRX_UNSUPPORTED, "unknown precision");
      return GPRX_UNSUPPORTED;
    }
    auto m = std::make_unique<ExactModel>();
    m->hx.resize(n * d);
    for (size_t i = 0; i < n; ++i)
      for (size_t q = 0; q < d; ++q) m->hx[i * d + q] = x[q * n + i];
    m->hy.assign(y, y + n);
    std::string why;
    int rc = parse_program(kernel, (int)d, &m->hp, &why);
    if (rc != GPRX_OK) {
      set_error(err, rc, why);
      return rc;
    }
    if (!log_theta) {
      set_error(err, GPRX_LENGTH_MISMATCH, "missing parameters");
      return GPRX_LENGTH_MISMATCH;
    }
    rc = validate_theta(m->hp, log_theta, p, err);
    if (rc != GPRX_OK) return rc;
    if (jitter) {
      if (!(jitter->initial >= 0.0) || !(jitter->multiplier >= 1.0) || jitter->max_retries < 0) {
        set_error(err, GPRX_INVALID_HYPERPARAMETER, "invalid jitter policy");
        return GPRX_INVALID_HYPERPARAMETER;
      }
      m->jit = *jitter;
    }
    m->n = (int)n;
    m->d = (int)d;
    m->dp = padded_dim((int)d);
    m->npad = (long)round_up(n, kTile);
    m->nt = (int)(m->npad / kTile);
    m->ldt = (long)((m->nt + 1) / 2) * kTile;
    m->precision = precision;
    m->theta.assign(log_theta, log_theta + p);
    GPRX_HIP_CHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    m->fstreams.init(m->stream);
    for (auto& e : m->ev) GPRX_HIP_CHECK(hipEventCreate(&e));
    hipStream_t s = m->stream;
    upload_points(m->x, x, (int)n, (int)d, m->npad, m->dp, s);
    {
      std::vector<double> yh(m->npad, 0.0);
      std::memcpy(yh.data(), y, sizeof(double) * n);
      m->y.alloc(m->npad);
      GPRX_HIP_CHECK(hipMemcpyAsync(m->y.p, yh.data(), sizeof(double) * m->npad, hipMemcpyHostToDevice, s));
      GPRX_HIP_CHECK(hipStreamSynchronize(s));
    }
    size_buffers(m.get());
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    *out = m.release();
    return GPRX_OK;
  });
}

int gprx_hip_exact_value_and_grad(gprx_hip_model* model, const double* log_theta, size_t p,
                                  double* nlml, double* grad, gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kExact || !log_theta || !nlml) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle or null output");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* m = static_cast<ExactModel*>(model);
    int rc = validate_theta(m->hp, log_theta, p, err);
    if (rc != GPRX_OK) return rc;
    std::vector<double> g(p, 0.0);
    double v = 0.0;
    rc = m->evaluate(log_theta, grad != nullptr, &v, grad ? g.data() : nullptr, err);
    if (rc != GPRX_OK) return rc;  // parameters unchanged; the factor is rebuilt on demand
    m->theta.assign(log_theta, log_theta + p);
    *nlml = v;
    if (grad) std::memcpy(grad, g.data(), sizeof(double) * p);
    return GPRX_OK;
  });
}

int gprx_hip_exact_hessian(gprx_hip_model* model, const double* log_theta, size_t p,
                           double* nlml, double* grad, double* hess, gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kExact || !log_theta || !nlml || !grad ||
        !hess) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle or null output");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* m = static_cast<ExactModel*>(model);
    int rc = validate_theta(m->hp, log_theta, p, err);
    if (rc != GPRX_OK) return rc;
    std::vector<double> g(p, 0.0), h(p * p, 0.0);
    double v = 0.0;
    rc = m->hessian(log_theta, &v, g.data(), h.data(), err);
    if (rc != GPRX_OK) return rc;
    m->theta.assign(log_theta, log_theta + p);
    *nlml = v;
    std::memcpy(grad, g.data(), sizeof(double) * p);
    std::memcpy(hess, h.data(), sizeof(double) * p * p);
    return GPRX_OK;
  });
}

// OnlineGpr::delete (src/gpr/online.rs:564-620): drops the point at row
// `index` and factors the model again (O(n^3); the reference downdates an
// LDL^T factor, which this model, holding L^-1 rather than L, does not have).
// The FP64 model's deletion in O(n^2) (the downdate above), then alpha =
// X^T (X y) and the NLML.  The factor keeps the jitter it was formed with,
// as the reference's downdated LDL^T does.
int exact_delete_downdate(ExactModel* m, int p, double* nlml, gprx_error* err) {
  hipStream_t s = m->stream;
  if (!m->factored || m->cur_form != GPRX_FP64) {
    double v0 = 0.0;
    const int rc = m->evaluate(m->theta.data(), false, &v0, nullptr, err);
    if (rc != GPRX_OK) return rc;
  }
  const int n = m->n, d = m->d;
  const long np = m->npad;
  const int mt = (int)(np - 1 - p);
  // the downdate's coefficients from column p of X, on the device
  if (mt > 0) {
    hipLaunchKernelGGL(k_delete_coeffs, dim3(1), dim3(kCoeffThreads), 0, s, m->X64.p, np, p, mt,
                       m->r.p, m->dlt.p, m->p1.p, m->q1.p);
    GPRX_HIP_CHECK(hipGetLastError());
  }
  // the downdated X into A64 (free between evaluations), then the buffers swap
  hipLaunchKernelGGL(k_delete_rows, dim3((unsigned)np), dim3(256), 0, s, m->X64.p, m->A64.p, np, p,
                     m->r.p, m->dlt.p, m->p1.p);
  GPRX_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_delete_logdiag, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s,
                     m->logdiag.p, m->q1.p, m->tv.p, p, np);
  GPRX_HIP_CHECK(hipGetLastError());
  GPRX_HIP_CHECK(hipMemcpyAsync(m->logdiag.p, m->tv.p, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
  std::swap(m->X64.p, m->A64.p);
  std::swap(m->X64.n, m->A64.n);
  // the points and targets without row p (rows keep their order), on the
  // device and in the host copies
  if (p < n - 1) {
    m->kcol.alloc((size_t)np * kTile);  // scratch: (n - 1 - p) (dp + 1) <= np kTile
    const long cnt = (long)(n - 1 - p) * (m->dp + 1);
    const unsigned blocks = (unsigned)std::min<long>((cnt + 255) / 256, 4096);
    hipLaunchKernelGGL(k_shift_rows_out, dim3(blocks), dim3(256), 0, s, m->x.p, np, m->dp, m->y.p, n,
                       p, m->kcol.p);
    hipLaunchKernelGGL(k_shift_rows_in, dim3(blocks), dim3(256), 0, s, m->x.p, np, m->dp, m->y.p, n, p,
                       m->kcol.p);
    GPRX_HIP_CHECK(hipGetLastError());
  } else {  // the last point: its row becomes padding
    hipLaunchKernelGGL(k_shift_rows_in, dim3(1), dim3(256), 0, s, m->x.p, np, m->dp, m->y.p, n, p,
                       m->kcol.p);
    GPRX_HIP_CHECK(hipGetLastError());
  }
  m->hx.erase(m->hx.begin() + (long)p * d, m->hx.begin() + (long)(p + 1) * d);
  m->hy.erase(m->hy.begin() + p);
  m->n = n - 1;
  // alpha = X^T (X y) and the NLML's sums
  tile_mv_lower<double>(m->X64.p, np, m->nt, kTmvRow, m->y.p, m->tv.p, nullptr, m->tpart.p, s);
  tile_mv_lower<double>(m->X64.p, np, m->nt, kTmvCol, m->tv.p, m->alpha.p, nullptr, m->tpart.p, s);
  dot_and_sum(m->y.p, m->alpha.p, m->logdiag.p, m->n, m->scal.p, s);
  GPRX_HIP_CHECK(hipMemcpyAsync(m->hscal.p, m->scal.p, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
  GPRX_HIP_CHECK(hipStreamSynchronize(s));
  m->z_valid = false;
  const double v = 0.5 * m->hscal.p[0] + m->hscal.p[1] + 0.5 * (double)m->n * kLog2Pi;
  if (nlml) *nlml = v;
  return GPRX_OK;
}

int exact_delete(ExactModel* m, size_t index, double* nlml, gprx_error* err) {
  if (m->precision == GPRX_FP64) return exact_delete_downdate(m, (int)index, nlml, err);
  // the mixed form factors the remaining points again
  const int n = m->n, d = m->d;
  std::vector<double> hx_new = m->hx;
  hx_new.erase(hx_new.begin() + (long)index * d, hx_new.begin() + (long)(index + 1) * d);
  std::vector<double> hx_old = m->hx, hy_old = m->hy;
  const long npad_old = m->npad;
  m->hx = hx_new;
  m->hy.erase(m->hy.begin() + (long)index);
  m->n = n - 1;
  regrow(m, (long)round_up((size_t)(n - 1), kTile));
  double v = 0.0;
  int rc = m->evaluate(m->theta.data(), false, &v, nullptr, err);
  if (rc != GPRX_OK) {  // back to the model before the deletion
    m->hx = hx_old;
    m->hy = hy_old;
    m->n = n;
    regrow(m, npad_old);
    gprx_error e2{};
    double v2 = 0.0;
    (void)m->evaluate(m->theta.data(), false, &v2, nullptr, &e2);
    return rc;
  }
  if (nlml) *nlml = v;
  return GPRX_OK;
}

int gprx_hip_exact_delete(gprx_hip_model* model, size_t index, double* nlml, gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kExact) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* m = static_cast<ExactModel*>(model);
    if (m->n <= 1) {
      set_error(err, GPRX_INSUFFICIENT_DATA, "deleting would leave fewer than one point");
      return GPRX_INSUFFICIENT_DATA;
    }
    if (index >= (size_t)m->n) {
      set_error(err, GPRX_LENGTH_MISMATCH, "no point at that index");
      return GPRX_LENGTH_MISMATCH;
    }
    return exact_delete(m, index, nlml, err);
  });
}

int gprx_hip_exact_insert(gprx_hip_model* model, const double* x_new, size_t d, double y_new,
                          double* nlml, gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kExact || !x_new) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle or null point");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* m = static_cast<ExactModel*>(model);
    if (d != (size_t)m->d) {
      set_error(err, GPRX_LENGTH_MISMATCH, "point dimension differs from training");
      return GPRX_LENGTH_MISMATCH;
    }
    if (m->n >= (1 << 20)) {
      set_error(err, GPRX_UNSUPPORTED, "exact GPR above 2^20 points");
      return GPRX_UNSUPPORTED;
    }
    return exact_insert(m, x_new, y_new, nlml, err);
  });
}

int gprx_hip_exact_export_factor(gprx_hip_model* model, double* l, size_t n, double* alpha,
                                 double* jitter, gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kExact || !l || !alpha || !jitter) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle or null output");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* m = static_cast<ExactModel*>(model);
    if (n != (size_t)m->n) {
      set_error(err, GPRX_LENGTH_MISMATCH, "n differs from the model's point count");
      return GPRX_LENGTH_MISMATCH;
    }
    // The FP64 evaluation leaves X64 = L^-1 (the factor itself is never
    // stored); L^T = L^-1 (Kc + s I) with Kc the kernel matrix and s the
    // noise variance plus the jitter the factor took, one triangular GEMM.
    const int saved = m->precision;
    m->precision = GPRX_FP64;
    double v = 0.0;
    const int rc = m->evaluate(m->theta.data(), false, &v, nullptr, err);
    m->precision = saved;
    if (rc != GPRX_OK) return rc;
    hipStream_t s = m->stream;
    const long np2 = m->npad * m->npad;
    const double factor_jitter = m->last_jitter;
    const double shift = std::exp(m->theta.back()) + factor_jitter;
    fill_cross<double>(m->dp, m->prog.p, m->pts(), m->pts(), m->A64.p, m->npad, s);
    DevBuf<double> lt;
    lt.alloc((size_t)np2);
    GPRX_HIP_CHECK(hipMemcpyAsync(lt.p, m->X64.p, sizeof(double) * np2, hipMemcpyDeviceToDevice, s));
    gemm_call<double>(false, false, m->X64.p, m->npad, m->A64.p, m->npad, lt.p, m->npad,
                      (int)m->npad, (int)m->npad, (int)m->npad, 1.0, shift, kKhiRow, 0,
                      kOrderColFastRev, s);
    // L = (L^T)^T into A64 (free until the next evaluation), then straight
    // into the caller's column-major l
    hipLaunchKernelGGL(k_transpose_lower, dim3(m->npad / 64, m->npad / 64), dim3(256), 0, s, lt.p,
                       m->A64.p, m->npad);
    GPRX_HIP_CHECK(hipGetLastError());
    GPRX_HIP_CHECK(hipMemcpy2DAsync(l, sizeof(double) * n, m->A64.p, sizeof(double) * m->npad,
                                    sizeof(double) * n, n, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(alpha, m->alpha.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    if (saved != GPRX_FP64) {  // the model predicts in its own form again
      const int rc2 = m->evaluate(m->theta.data(), false, &v, nullptr, err);
      if (rc2 != GPRX_OK) return rc2;
      // and gives back the FP64 factor's buffers (3 n_pad^2 doubles at most);
      // a later FP64 fallback allocates them again
      GPRX_HIP_CHECK(hipStreamSynchronize(s));
      for (DevBuf<double>* b : {&m->A64, &m->X64, &m->T64}) b->release();
    }
    *jitter = factor_jitter;
    return GPRX_OK;
  });
}

int gprx_hip_exact_predict(gprx_hip_model* model, const double* xs, size_t q, size_t d,
                           int32_t variance_kind, double* mean, double* var, gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kExact || !mean) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle or null output");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* m = static_cast<ExactModel*>(model);
    if (d != (size_t)m->d) {
      set_error(err, GPRX_LENGTH_MISMATCH, "query dimension differs from training");
      return GPRX_LENGTH_MISMATCH;
    }
    if (q == 0) return GPRX_OK;
    if (!xs || !all_finite(xs, q * d)) {
      set_error(err, GPRX_NON_FINITE_INPUT, "non-finite query");
      return GPRX_NON_FINITE_INPUT;
    }
    if (!m->factored) {
      double v;
      int rc = m->evaluate(m->theta.data(), false, &v, nullptr, err);
      if (rc != GPRX_OK) {
        if (err) err->stage = GPRX_STAGE_PREDICT;
        return rc;
      }
    } else {
      m->upload_program(m->theta.data(), m->last_jitter);
    }
    hipStream_t s = m->stream;
    const long qpad = (long)round_up(q, kTile);
    DevBuf<double> xq, Ks, kss, ss, mu;
    upload_points(xq, xs, (int)q, (int)d, qpad, m->dp, s);
    DevPoints pq{xq.p, qpad, (int)q};
    Ks.alloc((size_t)m->npad * qpad);
    kss.alloc(qpad);
    ss.alloc(qpad);
    mu.alloc(qpad);
    fill_cross<double>(m->dp, m->prog.p, m->pts(), pq, Ks.p, m->npad, s);
    gemv_t<double>(Ks.p, m->npad, (int)m->npad, (int)qpad, m->alpha.p, mu.p, 0, 0, 0, 0, s);
    const bool want_var = var != nullptr;
    if (want_var) {
      fill_diag(m->dp, m->prog.p, pq, kss.p, s);
      if (m->cur_form == GPRX_FP64) {
        DevBuf<double> V;
        V.alloc((size_t)m->npad * qpad);
        gemm_call<double>(false, false, m->X64.p, m->npad, Ks.p, m->npad, V.p, m->npad,
                          (int)m->npad, (int)qpad, (int)m->npad, 1.0, 0.0, kKhiRow, 0,
                          kOrderColFastRev, s);
        col_sumsq<double>(V.p, m->npad, (int)m->npad, (int)qpad, ss.p, s);
        GPRX_HIP_CHECK(hipStreamSynchronize(s));
      } else {
        DevBuf<float> K32, V;
        K32.alloc((size_t)m->npad * qpad);
        V.alloc((size_t)m->npad * qpad);
        fill_cross<float>(m->dp, m->prog.p, m->pts(), pq, K32.p, m->npad, s);
        gemm_call<float>(false, false, m->X32.p, m->npad, K32.p, m->npad, V.p, m->npad,
                         (int)m->npad, (int)qpad, (int)m->npad, 1.0, 0.0, kKhiRow, 0,
                         kOrderColFastRev, s);
        col_sumsq<float>(V.p, m->npad, (int)m->npad, (int)qpad, ss.p, s);
        GPRX_HIP_CHECK(hipStreamSynchronize(s));
      }
    }
    GPRX_HIP_CHECK(hipMemcpyAsync(mean, mu.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
    std::vector<double> hk(q), hs(q);
    if (want_var) {
      GPRX_HIP_CHECK(hipMemcpyAsync(hk.data(), kss.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
      GPRX_HIP_CHECK(hipMemcpyAsync(hs.data(), ss.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
    }
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    if (want_var) {
      const double noise = std::exp(m->theta[m->hp.n_params]);
      for (size_t j = 0; j < q; ++j) {
        double lat = hk[j] - hs[j];
        if (lat < 0.0) lat = 0.0;  // src/gpr/shared.rs:798-806
        var[j] = variance_kind == GPRX_VARIANCE_LATENT ? lat : lat + noise;
      }
    }
    return GPRX_OK;
  });
}

int gprx_hip_exact_predict_noalloc(gprx_hip_model* model, const double* xs, size_t q, size_t d,
                                   int32_t variance_kind, double* mean, double* var,
                                   gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kExact || !mean) {
      set_error(err, GPRX_LENGTH_MISMATCH, "bad handle or null output");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* m = static_cast<ExactModel*>(model);
    if (d != (size_t)m->d) {
      set_error(err, GPRX_LENGTH_MISMATCH, "query dimension differs from training");
      return GPRX_LENGTH_MISMATCH;
    }
    if (m->npad > kNaRows || m->dp > kNaDims) {
      set_error(err, GPRX_UNSUPPORTED,
                "the no-allocation predict takes models of at most 16384 (padded) points");
      return GPRX_UNSUPPORTED;
    }
    if (!m->factored) {
      set_error(err, GPRX_INVALID_CONFIG,
                "the no-allocation predict needs a factored model (evaluate it first)");
      return GPRX_INVALID_CONFIG;
    }
    if (q == 0) return GPRX_OK;
    if (!xs || !all_finite(xs, q * d)) {
      set_error(err, GPRX_NON_FINITE_INPUT, "non-finite query");
      return GPRX_NON_FINITE_INPUT;
    }
    int dev = 0;
    double* sp = nullptr;
    const int rs = na_scratch(&dev, &sp, err);
    if (rs != GPRX_OK) return rs;
    std::lock_guard<std::mutex> lock(g_na_call_mu[dev]);
    hipStream_t s = m->stream;
    m->upload_program(m->theta.data(), m->last_jitter);
    const long np = m->npad;
    double* xq = sp + kNaOffX;
    double* Ks = sp + kNaOffK;
    double* V = sp + kNaOffV;
    float* K32 = reinterpret_cast<float*>(sp + kNaOffV);
    float* V32 = K32 + np * kNaCols;
    double* kss = sp + kNaOffVec;
    double* ss = kss + kNaCols;
    double* mu = ss + kNaCols;
    double* omean = mu + kNaCols;
    double* ovar = omean + kNaCols;
    const bool want_var = var != nullptr;
    const double noise = std::exp(m->theta[m->hp.n_params]);
    for (size_t j0 = 0; j0 < q; j0 += (size_t)kNaCols) {
      const int qc = (int)std::min<size_t>((size_t)kNaCols, q - j0);
      // the chunk's points, column-major qc x d into kNaCols x dp with the
      // padded dimensions zero
      zero<double>(xq, (size_t)kNaCols * m->dp, s);
      GPRX_HIP_CHECK(hipMemcpy2DAsync(xq, sizeof(double) * kNaCols, xs + j0, sizeof(double) * q,
                                      sizeof(double) * qc, d, hipMemcpyHostToDevice, s));
      DevPoints pq{xq, kNaCols, qc};
      fill_cross<double>(m->dp, m->prog.p, m->pts(), pq, Ks, np, s);
      gemv_t<double>(Ks, np, (int)np, (int)kNaCols, m->alpha.p, mu, 0, 0, 0, 0, s);
      if (want_var) {
        fill_diag(m->dp, m->prog.p, pq, kss, s);
        if (m->cur_form == GPRX_FP64) {
          gemm_call<double>(false, false, m->X64.p, np, Ks, np, V, np, (int)np, (int)kNaCols,
                            (int)np, 1.0, 0.0, kKhiRow, 0, kOrderColFastRev, s, false);
          col_sumsq<double>(V, np, (int)np, (int)kNaCols, ss, s);
        } else {
          fill_cross<float>(m->dp, m->prog.p, m->pts(), pq, K32, np, s);
          gemm_call<float>(false, false, m->X32.p, np, K32, np, V32, np, (int)np, (int)kNaCols,
                           (int)np, 1.0, 0.0, kKhiRow, 0, kOrderColFastRev, s, false);
          col_sumsq<float>(V32, np, (int)np, (int)kNaCols, ss, s);
        }
      }
      hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(kNaCols), 0, s, mu, kss, ss, qc, want_var ? 1 : 0,
                         variance_kind == GPRX_VARIANCE_LATENT ? 1 : 0, noise, omean, ovar);
      GPRX_HIP_CHECK(hipGetLastError());
      GPRX_HIP_CHECK(hipMemcpyAsync(mean + j0, omean, sizeof(double) * qc, hipMemcpyDeviceToHost, s));
      if (want_var)
        GPRX_HIP_CHECK(hipMemcpyAsync(var + j0, ovar, sizeof(double) * qc, hipMemcpyDeviceToHost, s));
    }
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    return GPRX_OK;
  });
}

int gprx_hip_get_params(const gprx_hip_model* model, double* log_theta, size_t p) {
  if (!model || !log_theta) return GPRX_LENGTH_MISMATCH;
  if (model->kind != gprx_hip_model::kExact) return GPRX_UNSUPPORTED;
  auto* m = static_cast<const ExactModel*>(model);
  if (p != m->theta.size()) return GPRX_LENGTH_MISMATCH;
  std::memcpy(log_theta, m->theta.data(), sizeof(double) * p);
  return GPRX_OK;
}

double gprx_hip_last_jitter(const gprx_hip_model* model) {
  if (!model || model->kind != gprx_hip_model::kExact) return 0.0;
  return static_cast<const ExactModel*>(model)->last_jitter;
}

int gprx_hip_last_fallback(const gprx_hip_model* model) {
  if (!model || model->kind != gprx_hip_model::kExact) return 0;
  return static_cast<const ExactModel*>(model)->fallback;
}

int gprx_hip_set_mixed_jitter_bound(gprx_hip_model* model, double bound, gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kExact) {
      set_error(err, GPRX_UNSUPPORTED, "not an exact model");
      return GPRX_UNSUPPORTED;
    }
    if (!(bound >= 0.0)) {
      set_error(err, GPRX_INVALID_HYPERPARAMETER, "mixed jitter bound must be >= 0");
      return GPRX_INVALID_HYPERPARAMETER;
    }
    static_cast<ExactModel*>(model)->mixed_jitter_bound = bound;
    return GPRX_OK;
  });
}

int gprx_hip_last_refine_steps(const gprx_hip_model* model) {
  if (!model || model->kind != gprx_hip_model::kExact) return 0;
  return static_cast<const ExactModel*>(model)->refine_steps;
}

void gprx_hip_destroy(gprx_hip_model* model) { delete model; }

double gprx_hip_last_eval_ms(const gprx_hip_model* model) { return model ? model->last_eval_ms : 0.0; }
double gprx_hip_last_dense_ms(const gprx_hip_model* model) {
  return model ? model->last_dense_ms : 0.0;
}

double gprx_hip_last_phase_ms(const gprx_hip_model* model, int32_t phase) {
  if (!model || model->kind != gprx_hip_model::kExact || phase < 0 || phase > 4) return 0.0;
  return static_cast<const ExactModel*>(model)->phase_ms[phase];
}

}  // extern "C"
