// Model handles behind the C ABI.
#pragma once

#include "common.hpp"
#include "dense.hpp"
#include "fill.hpp"

struct gprx_hip_model {
  enum Kind { kExact = 1, kSgpr = 2, kSvgp = 3, kBatch = 4 };
  explicit gprx_hip_model(Kind k) : kind(k) {}
  virtual ~gprx_hip_model() = default;
  Kind kind;
  double last_eval_ms = 0.0;
  double last_dense_ms = 0.0;
};

namespace gprx {

// Small on-device vector helpers shared by the models (vec.hip).
// out[0] = sum_{i<n} a_i b_i, out[1] = sum_{i<n} c_i  (one block, fixed order)
void dot_and_sum(const double* a, const double* b, const double* c, int n, double* out,
                 hipStream_t s);
// out[0] = sum_{i<n} (alpha_i^2 - M_ii) over the diagonal of M (storage T)
template <typename T>
void noise_trace(const double* alpha, const T* M, long ld, int n, double* out, hipStream_t s);
// r = y - (p1 + p2) for i < n (zero beyond); out[0] = max|r|, out[1] = max|w|,
// out[2] = max|y|, out[3] = max(rowsum) of the absolute row sums in q1 + q2
// (q1 null: out[3] untouched; p2 / q2 null: p1 / q1 alone).
void residual_norms(const double* y, const double* p1, const double* p2, const double* w,
                    const double* q1, const double* q2, double* r, int n, double* out,
                    hipStream_t s);
// w += d for i < n
void axpy(double* w, const double* d, int n, hipStream_t s);
// dst (T) = src (U) elementwise, count entries.
template <typename T, typename U>
void convert(T* dst, const U* src, size_t count, hipStream_t s);

// Host helpers.
bool all_finite(const double* p, size_t n);

// Tile-GEMM call in one line.
template <typename T>
inline void gemm_call(bool ta, bool tb, const T* A, long lda, const T* B, long ldb, T* C, long ldc,
                      int M, int N, int K, double alpha, double beta, int krange, int lower,
                      int order, hipStream_t s, bool scheduled = true) {
  GemmArgs<T> g{};
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.C = C;
  g.ldc = ldc;
  g.M = M;
  g.N = N;
  g.K = K;
  g.alpha = alpha;
  g.beta = beta;
  g.krange = krange;
  g.lower = lower;
  g.order = order;
  g.batch = 1;
  if (scheduled)
    gemm<T>(ta, tb, g, s);
  else
    gemm_unscheduled<T>(ta, tb, g, s);
}

// Streams of the recursive factor.  `main` carries the recursion and its
// latency-bound chain of small launches; each node's product Tmp = L21 X11,
// read only by the node's last step, runs on `bulk` meanwhile.  `bulk` is
// one stream per device shared by every model, restricted to all CUs but
// kReservedCus (4 on each of the 8 XCDs), so the chain's kernels always
// find free CUs while the bulk product fills the rest.  Per recursion
// depth k, ev[2k] hands L21 to bulk and ev[2k+1] returns Tmp.  `ahead`,
// restricted like bulk, takes the two top depths' lookahead (the first
// columns of L21 and their share of the update, launched once a block's
// first quarter is factored); aev[2k] hands X11's first quarter to it and
// aev[2k+1] returns the update.  All three streams keep the default
// priority.
constexpr int kMaxDepth = 24;
// Measured on gfx950 at n = 16384 (factor + inverse, f64 / f32): 52.3 /
// 29.1 ms with 32 reserved CUs, 53.4 / 29.6 with 16, 54.7 / 29.9 with 48.
constexpr int kReservedCus = 32;
// Shared bulk stream of the current device (created on first use).
hipStream_t bulk_stream();
// Shared lookahead stream of the current device (created on first use).
hipStream_t ahead_stream();
struct FactorStreams {
  hipStream_t main = nullptr, bulk = nullptr, ahead = nullptr;  // bulk, ahead shared
  hipEvent_t ev[2 * kMaxDepth] = {};
  hipEvent_t aev[4] = {};  // lookahead hand-over per depth 0, 1
  void init(hipStream_t s) {
    main = s;
    bulk = bulk_stream();
    ahead = ahead_stream();
    for (auto& e : ev) GPRX_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : aev) GPRX_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  // Every bulk product is waited for on main before the factor returns, so
  // once main is synchronized no bulk work of this model is pending.
  void destroy() {
    for (auto& e : ev)
      if (e) {
        (void)hipEventDestroy(e);
        e = nullptr;
      }
    for (auto& e : aev)
      if (e) {
        (void)hipEventDestroy(e);
        e = nullptr;
      }
    bulk = nullptr;
    ahead = nullptr;
  }
};

// Columns (in 128-tiles) of the Tmp workspace of factor_inverse for nt
// tiles: one block per recursion depth, ceil(nt/2) tiles of rows (ldt =
// ceil(nt/2) * 128), since a depth's Tmp must outlive the deeper ones.
inline long factor_tmp_cols(int nt) {
  long cols = 0;
  for (int s = nt; s > 1; s = s - s / 2) cols += s / 2;
  return cols > 0 ? cols : 1;
}

// Recursive blocked Cholesky + triangular inverse of the lower part of the
// nt*128 square A (ld), writing X = L^-1 (lower) into X and the log of every
// pivot into logdiag.  A's lower triangle is consumed; Tmp holds ldt x
// factor_tmp_cols(nt) * 128 (ldt = ceil(nt/2) * 128).
template <typename T>
void factor_inverse(T* A, T* X, T* Tmp, long ld, long ldt, int nt, double* logdiag, int* info,
                    const FactorStreams& fs);

}  // namespace gprx

namespace gprx {
// Dense m x m helpers (f64, column-major, full storage).
// out = a * A + b * B over rows x cols (out may alias A or B).
void mat_combine(double* out, long ldo, int rows, int cols, double a, const double* A, long lda,
                 double b, const double* B, long ldb, hipStream_t s);
// Copies the strict lower triangle onto the upper one.
void symmetrize_lower(double* M, long ld, int n, hipStream_t s);
// out[0] = sum_ij A_ij B_ij over rows x cols (one block, fixed order).
void frob_dot(const double* A, long lda, const double* B, long ldb, int rows, int cols, double* out,
              hipStream_t s);
// W += alpha * u v^T over rows x cols.
void outer_add(double* W, long ld, int rows, int cols, double alpha, const double* u,
               const double* v, hipStream_t s);
// out_i = a * y_i + b * v_i for i < n, zero for n <= i < npad.
void vec_combine(double* out, int n, int npad, double a, const double* y, double b, const double* v,
                 hipStream_t s);
// Sums `parts` lower-triangle partial matrices (storage T, each ld x ld,
// consecutive) into out (f64), fixed order.
template <typename T>
void sum_partials_lower(const T* parts, int nparts, long ld, int n, double* out, hipStream_t s);
}  // namespace gprx

namespace gprx {
// M_ii += v for i < n.
void add_diag(double* M, long ld, int n, double v, hipStream_t s);
// out[0] = sum_{i<n} M_ii (one block).
void diag_sum(const double* M, long ld, int n, double* out, hipStream_t s);
// H = tril_half(W): strict lower of W, half its diagonal, zero above (W of storage T).
template <typename T>
void tril_half(const T* W, long ld, int n, double* H, hipStream_t s);
// W = -1/2 (T2 + T2^T) over n x n.
void neg_sym(const double* T2, long ld, int n, double* W, hipStream_t s);
}  // namespace gprx
