// Small vector kernels and the recursive factor + inverse driver.
#include <map>
#include <mutex>

#include "model.hpp"

namespace gprx {

bool all_finite(const double* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

// One 1024-thread block; wavefront-shuffle reductions in a fixed order.
// out[0] = a.b (y'alpha), out[1] = sum c (the log pivots: log |A| / 2).
__global__ __launch_bounds__(1024) void k_dot_and_sum(const double* a, const double* b,
                                                      const double* c, int n, double* out) {
  __shared__ double r0[16], r1[16];
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    s0 += a[i] * b[i];
    s1 += c[i];
  }
  s0 = block_sum<1024>(s0, r0);
  s1 = block_sum<1024>(s1, r1);
  if (threadIdx.x == 0) {
    out[0] = s0;
    out[1] = s1;
  }
}

void dot_and_sum(const double* a, const double* b, const double* c, int n, double* out,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_dot_and_sum, dim3(1), dim3(1024), 0, s, a, b, c, n, out);
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T>
__global__ __launch_bounds__(1024) void k_noise_trace(const double* alpha, const T* M, long ld,
                                                      int n, double* out) {
  __shared__ double r0[16];
  double s0 = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s0 += alpha[i] * alpha[i] - (double)M[(long)i * (ld + 1)];
  s0 = block_sum<1024>(s0, r0);
  if (threadIdx.x == 0) out[0] = s0;
}

template <typename T>
void noise_trace(const double* alpha, const T* M, long ld, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL((k_noise_trace<T>), dim3(1), dim3(1024), 0, s, alpha, M, ld, n, out);
  GPRX_HIP_CHECK(hipGetLastError());
}
template void noise_trace<double>(const double*, const double*, long, int, double*, hipStream_t);
template void noise_trace<float>(const double*, const float*, long, int, double*, hipStream_t);

__global__ __launch_bounds__(1024) void k_residual_norms(const double* y, const double* p1,
                                                         const double* p2, const double* w,
                                                         const double* q1, const double* q2,
                                                         double* r, int n, double* out) {
  __shared__ double m0[16], m1[16], m2[16], m3[16];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    const double ri = y[i] - (p2 ? p1[i] + p2[i] : p1[i]);
    r[i] = ri;
    a0 = fmax(a0, fabs(ri));
    a1 = fmax(a1, fabs(w[i]));
    a2 = fmax(a2, fabs(y[i]));
    if (q1) a3 = fmax(a3, q2 ? q1[i] + q2[i] : q1[i]);
  }
  a0 = block_max<1024>(a0, m0);
  a1 = block_max<1024>(a1, m1);
  a2 = block_max<1024>(a2, m2);
  a3 = block_max<1024>(a3, m3);
  if (threadIdx.x == 0) {
    out[0] = a0;
    out[1] = a1;
    out[2] = a2;
    if (q1) out[3] = a3;
  }
}

void residual_norms(const double* y, const double* p1, const double* p2, const double* w,
                    const double* q1, const double* q2, double* r, int n, double* out,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_residual_norms, dim3(1), dim3(1024), 0, s, y, p1, p2, w, q1, q2, r, n, out);
  GPRX_HIP_CHECK(hipGetLastError());
}

__global__ void k_axpy(double* w, const double* d, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) w[i] += d[i];
}

void axpy(double* w, const double* d, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, s, w, d, n);
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T, typename U>
__global__ void k_convert(T* dst, const U* src, size_t count) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += stride) dst[i] = (T)src[i];
}

template <typename T, typename U>
void convert(T* dst, const U* src, size_t count, hipStream_t s) {
  if (count == 0) return;
  size_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((k_convert<T, U>), dim3((unsigned)blocks), dim3(256), 0, s, dst, src, count);
  GPRX_HIP_CHECK(hipGetLastError());
}
template void convert<float, double>(float*, const double*, size_t, hipStream_t);
template void convert<double, float>(double*, const float*, size_t, hipStream_t);
template void convert<double, double>(double*, const double*, size_t, hipStream_t);

// ---------------------------------------------------------------------------
// Recursive Cholesky + inverse.  For the block [A11 . ; A21 A22] of s tiles
// (split s1 = s/2, s2 = s - s1):
//   (L11, X11) = rec(A11)
//   L21  = A21 X11^T                 (stored in X's 21 block meanwhile)
//   A22 -= L21 L21^T                 (lower tiles)
//   Tmp  = L21 X11
//   (L22, X22) = rec(A22)
//   X21  = -X22 Tmp
// Only X = L^-1 leaves the recursion (the exact path needs A^-1 = X^T X,
// alpha = X^T X y and predictions through X), so L is never stored.
// Flops: n^3/3 for the factor plus n^3/3 for the inverse, the large products
// sitting at the top of the recursion.
//
// Overlap: in a block of at least kOverlapMin tiles, Tmp (read only by the
// block's last step) is formed on the bulk stream, which leaves four CUs of
// every XCD free, while main applies the update and runs rec(A22), a chain
// of launches too small to fill the chip.  Measured on gfx950 at n = 16384
// (factor + inverse, f32 / f64): 38.3 / 66.0 ms, against 41.6 / 71.2 ms on
// one stream.  Splitting the update as well (its far rows on bulk, so
// rec(A22) starts sooner) measured slower, 39.7 / 69.8 ms: the split
// launches have worse tile tails and queue ahead of Tmp on bulk.
// ---------------------------------------------------------------------------
constexpr int kOverlapMin = 8;

// A stream per (device, slot) restricted to all CUs but kReservedCus.  A
// process keeps two of these: each holds a hardware queue, and with a third
// one beside two models' own streams every kernel of the second model ran
// 5-150% slower (measured; most likely the queues beyond those the device
// maps at once being time-sliced).
static hipStream_t masked_stream(int slot) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, hipStream_t> streams;
  int dev = 0;
  GPRX_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto it = streams.find({dev, slot});
  if (it != streams.end()) return it->second;
  hipDeviceProp_t prop;
  GPRX_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  const int ncu = prop.multiProcessorCount;
  hipStream_t st = nullptr;
  // CU mask bit b selects a CU on XCD b mod 8, so bits 0..31 are four CUs
  // of every XCD.  Other CU counts get an unrestricted stream.
  if (ncu == 256) {
    uint32_t mask[8];
    for (auto& w : mask) w = 0xffffffffu;
    for (int b = 0; b < kReservedCus; ++b) mask[b / 32] &= ~(1u << (b % 32));
    GPRX_HIP_CHECK(hipExtStreamCreateWithCUMask(&st, 8, mask));
  } else {
    GPRX_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  }
  streams[{dev, slot}] = st;  // kept until the process ends
  return st;
}

hipStream_t bulk_stream() { return masked_stream(0); }
hipStream_t ahead_stream() { return masked_stream(1); }

// Lookahead.  A block N of s >= kAheadMin tiles at recursion depth <=
// kAheadMaxDepth splits L21 by columns at q = s1 / 2: L21's first q columns
// need only X of N's first quarter, so they and their share of the A22
// update (A22 -= L21a L21a^T) are launched on the ahead stream as soon as
// that quarter is factored, and fill the chip while main works through the
// rest of rec(A11).  After rec(A11) main forms the other columns, waits for
// the ahead stream and applies the rest of the update.  Only the two top
// depths take part: a deeper block inside a first quarter's sibling would
// queue behind its ancestor's larger lookahead on the one ahead stream.
// Measured on gfx950 at n = 16384 (factor + inverse, f64 / f32): 53.9 /
// 30.3 ms against 54.4 / 30.9 ms without; the ahead stream restricted to
// half the CUs gave 54.3 / 30.2 ms, to a quarter 69.3 / 37.3 ms (main then
// waits for it).
constexpr int kAheadMaxDepth = 1;
constexpr int kAheadMin = 16;

struct AheadNode {
  int o, s, depth;
};

template <typename T>
static void launch_ahead(T* A, T* X, long ld, const AheadNode& p, const FactorStreams& fs) {
  auto at = [&](T* base, int r, int c) { return base + (long)r * kTile + (long)c * kTile * ld; };
  const int s1 = p.s / 2, q = s1 / 2, s2 = p.s - s1, o2 = p.o + s1;
  const hipEvent_t* ev = fs.aev + 2 * p.depth;
  GPRX_HIP_CHECK(hipEventRecord(ev[0], fs.main));
  GPRX_HIP_CHECK(hipStreamWaitEvent(fs.ahead, ev[0], 0));
  // L21[:, :q] = A21[:, :q] X11[:q, :q]^T
  gemm_call<T>(false, true, at(A, o2, p.o), ld, at(X, p.o, p.o), ld, at(X, o2, p.o), ld, s2 * kTile,
               q * kTile, q * kTile, 1.0, 0.0, kKhiCol, 0, kOrderRowFastRev, fs.ahead);
  // A22 -= L21[:, :q] L21[:, :q]^T
  gemm_call<T>(false, true, at(X, o2, p.o), ld, at(X, o2, p.o), ld, at(A, o2, o2), ld, s2 * kTile,
               s2 * kTile, q * kTile, -1.0, 1.0, 0, 1, kOrderRowFast, fs.ahead);
  GPRX_HIP_CHECK(hipEventRecord(ev[1], fs.ahead));
}

template <typename T>
static void rec(T* A, T* X, T* Tmp, long ld, long ldt, int o, int s, int depth, const long* toff,
                double* logdiag, int* info, const FactorStreams& fs, const AheadNode* parent) {
  hipStream_t st = fs.main;
  auto at = [&](T* base, int r, int c) { return base + (long)r * kTile + (long)c * kTile * ld; };
  if (s == 1) {
    potrf_diag<T>(at(A, o, o), ld, at(X, o, o), ld, logdiag + (long)o * kTile, info, 1, 0, 0, 0, st);
    return;
  }
  if (depth >= kMaxDepth) throw DeviceError("factor recursion deeper than supported");
  const int s1 = s / 2, s2 = s - s1;
  const int o2 = o + s1;
  const bool ahead = fs.ahead && depth <= kAheadMaxDepth && s >= kAheadMin;
  const AheadNode self{o, s, depth};
  rec<T>(A, X, Tmp, ld, ldt, o, s1, depth + 1, toff, logdiag, info, fs, ahead ? &self : nullptr);
  // this block's first half is the parent's first quarter
  if (parent) launch_ahead<T>(A, X, ld, *parent, fs);
  const int q = ahead ? s1 / 2 : 0, sq = s1 - q;
  // L21 = A21 X11^T (X11 lower, stored N x K: k <= j); with lookahead only
  // its columns q.. here: X11[q:, :q] is full, X11[q:, q:] lower
  if (ahead) {
    gemm_call<T>(false, true, at(A, o2, o), ld, at(X, o + q, o), ld, at(X, o2, o + q), ld,
                 s2 * kTile, sq * kTile, q * kTile, 1.0, 0.0, 0, 0, kOrderRowFast, st);
    gemm_call<T>(false, true, at(A, o2, o + q), ld, at(X, o + q, o + q), ld, at(X, o2, o + q), ld,
                 s2 * kTile, sq * kTile, sq * kTile, 1.0, 1.0, kKhiCol, 0, kOrderRowFastRev, st);
    GPRX_HIP_CHECK(hipStreamWaitEvent(st, fs.aev[2 * depth + 1], 0));
  } else {
    gemm_call<T>(false, true, at(A, o2, o), ld, at(X, o, o), ld, at(X, o2, o), ld, s2 * kTile,
                 s1 * kTile, s1 * kTile, 1.0, 0.0, kKhiCol, 0, kOrderRowFastRev, st);
  }
  // Tmp = L21 X11 (X11 lower K x N: k >= j), this depth's block
  const bool overlap = fs.bulk && s >= kOverlapMin;
  const hipEvent_t* ev = fs.ev + 2 * depth;
  hipStream_t ts = st;
  if (overlap) {
    GPRX_HIP_CHECK(hipEventRecord(ev[0], st));
    GPRX_HIP_CHECK(hipStreamWaitEvent(fs.bulk, ev[0], 0));
    ts = fs.bulk;
  }
  T* Tk = Tmp + toff[depth] * kTile * ldt;
  gemm_call<T>(false, false, at(X, o2, o), ld, at(X, o, o), ld, Tk, ldt, s2 * kTile, s1 * kTile,
               s1 * kTile, 1.0, 0.0, kKloCol, 0, kOrderRowFast, ts);
  if (overlap) GPRX_HIP_CHECK(hipEventRecord(ev[1], fs.bulk));
  // A22 -= L21 L21^T (with lookahead the columns q.. only)
  gemm_call<T>(false, true, at(X, o2, o + q), ld, at(X, o2, o + q), ld, at(A, o2, o2), ld,
               s2 * kTile, s2 * kTile, sq * kTile, -1.0, 1.0, 0, 1, kOrderRowFast, st);
  rec<T>(A, X, Tmp, ld, ldt, o2, s2, depth + 1, toff, logdiag, info, fs, nullptr);
  // X21 = -X22 Tmp (X22 lower: k <= i), over L21's storage
  if (overlap) GPRX_HIP_CHECK(hipStreamWaitEvent(st, ev[1], 0));
  gemm_call<T>(false, false, at(X, o2, o2), ld, Tk, ldt, at(X, o2, o), ld, s2 * kTile, s1 * kTile,
               s2 * kTile, -1.0, 0.0, kKhiRow, 0, kOrderColFastRev, st);
}

template <typename T>
void factor_inverse(T* A, T* X, T* Tmp, long ld, long ldt, int nt, double* logdiag, int* info,
                    const FactorStreams& fs) {
  // Tmp column offset (tiles) of each depth, along the widest path
  long toff[kMaxDepth + 1] = {0};
  int depth = 0;
  for (int s = nt; s > 1 && depth < kMaxDepth; s = s - s / 2, ++depth) toff[depth + 1] = toff[depth] + s / 2;
  rec<T>(A, X, Tmp, ld, ldt, 0, nt, 0, toff, logdiag, info, fs, nullptr);
}

template void factor_inverse<double>(double*, double*, double*, long, long, int, double*, int*,
                                     const FactorStreams&);
template void factor_inverse<float>(float*, float*, float*, long, long, int, double*, int*,
                                    const FactorStreams&);

}  // namespace gprx

namespace gprx {

__global__ void k_mat_combine(double* out, long ldo, int rows, int cols, double a, const double* A,
                              long lda, double b, const double* B, long ldb) {
  const long total = (long)rows * cols;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % rows), j = (int)(e / rows);
    const double va = A ? A[i + j * lda] : 0.0;
    const double vb = B ? B[i + j * ldb] : 0.0;
    out[i + j * ldo] = a * va + b * vb;
  }
}

void mat_combine(double* out, long ldo, int rows, int cols, double a, const double* A, long lda,
                 double b, const double* B, long ldb, hipStream_t s) {
  const long total = (long)rows * cols;
  const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_mat_combine, dim3(blocks), dim3(256), 0, s, out, ldo, rows, cols, a, A, lda,
                     b, B, ldb);
  GPRX_HIP_CHECK(hipGetLastError());
}

__global__ void k_symmetrize(double* M, long ld, int n) {
  const long total = (long)n * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i < j) M[i + j * ld] = M[j + i * ld];
  }
}

void symmetrize_lower(double* M, long ld, int n, hipStream_t s) {
  const long total = (long)n * n;
  const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_symmetrize, dim3(blocks), dim3(256), 0, s, M, ld, n);
  GPRX_HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(1024) void k_frob_dot(const double* A, long lda, const double* B,
                                                   long ldb, int rows, int cols, double* out) {
  __shared__ double red[16];
  double acc = 0.0;
  const long total = (long)rows * cols;
  for (long e = threadIdx.x; e < total; e += 1024) {
    const int i = (int)(e % rows), j = (int)(e / rows);
    acc += A[i + j * lda] * B[i + j * ldb];
  }
  acc = block_sum<1024>(acc, red);
  if (threadIdx.x == 0) out[0] = acc;
}

void frob_dot(const double* A, long lda, const double* B, long ldb, int rows, int cols, double* out,
              hipStream_t s) {
  hipLaunchKernelGGL(k_frob_dot, dim3(1), dim3(1024), 0, s, A, lda, B, ldb, rows, cols, out);
  GPRX_HIP_CHECK(hipGetLastError());
}

__global__ void k_outer_add(double* W, long ld, int rows, int cols, double alpha, const double* u,
                            const double* v) {
  const long total = (long)rows * cols;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % rows), j = (int)(e / rows);
    W[i + j * ld] += alpha * u[i] * v[j];
  }
}

void outer_add(double* W, long ld, int rows, int cols, double alpha, const double* u,
               const double* v, hipStream_t s) {
  const long total = (long)rows * cols;
  const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_outer_add, dim3(blocks), dim3(256), 0, s, W, ld, rows, cols, alpha, u, v);
  GPRX_HIP_CHECK(hipGetLastError());
}

__global__ void k_vec_combine(double* out, int n, int npad, double a, const double* y, double b,
                              const double* v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= npad) return;
  out[i] = i < n ? a * y[i] + b * v[i] : 0.0;
}

void vec_combine(double* out, int n, int npad, double a, const double* y, double b, const double* v,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_vec_combine, dim3((npad + 255) / 256), dim3(256), 0, s, out, n, npad, a, y,
                     b, v);
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T>
__global__ void k_sum_partials_lower(const T* parts, int nparts, long ld, int n, double* out) {
  const long total = (long)n * n;
  const long stride = ld * ld;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i < j) continue;
    double acc = 0.0;
    for (int p = 0; p < nparts; ++p) acc += (double)parts[p * stride + i + j * ld];
    out[i + j * ld] = acc;
  }
}

template <typename T>
void sum_partials_lower(const T* parts, int nparts, long ld, int n, double* out, hipStream_t s) {
  const long total = (long)n * n;
  const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL((k_sum_partials_lower<T>), dim3(blocks), dim3(256), 0, s, parts, nparts, ld,
                     n, out);
  GPRX_HIP_CHECK(hipGetLastError());
}
template void sum_partials_lower<double>(const double*, int, long, int, double*, hipStream_t);
template void sum_partials_lower<float>(const float*, int, long, int, double*, hipStream_t);

}  // namespace gprx

namespace gprx {

__global__ void k_add_diag(double* M, long ld, int n, double v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) M[(long)i * (ld + 1)] += v;
}

void add_diag(double* M, long ld, int n, double v, hipStream_t s) {
  hipLaunchKernelGGL(k_add_diag, dim3((n + 255) / 256), dim3(256), 0, s, M, ld, n, v);
  GPRX_HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(1024) void k_diag_sum(const double* M, long ld, int n, double* out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) acc += M[(long)i * (ld + 1)];
  acc = block_sum<1024>(acc, red);
  if (threadIdx.x == 0) out[0] = acc;
}

void diag_sum(const double* M, long ld, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_sum, dim3(1), dim3(1024), 0, s, M, ld, n, out);
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T>
__global__ void k_tril_half(const T* W, long ld, int n, double* H) {
  const long total = (long)n * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    const double v = (double)W[i + j * ld];
    H[i + j * ld] = i > j ? v : (i == j ? 0.5 * v : 0.0);
  }
}

template <typename T>
void tril_half(const T* W, long ld, int n, double* H, hipStream_t s) {
  const long total = (long)n * n;
  const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL((k_tril_half<T>), dim3(blocks), dim3(256), 0, s, W, ld, n, H);
  GPRX_HIP_CHECK(hipGetLastError());
}
template void tril_half<double>(const double*, long, int, double*, hipStream_t);
template void tril_half<float>(const float*, long, int, double*, hipStream_t);

__global__ void k_neg_sym(const double* T2, long ld, int n, double* W) {
  const long total = (long)n * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    W[i + j * ld] = -0.5 * (T2[i + j * ld] + T2[j + i * ld]);
  }
}

void neg_sym(const double* T2, long ld, int n, double* W, hipStream_t s) {
  const long total = (long)n * n;
  const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_neg_sym, dim3(blocks), dim3(256), 0, s, T2, ld, n, W);
  GPRX_HIP_CHECK(hipGetLastError());
}

}  // namespace gprx
