// Kernel programs: host parsing into sum-of-products form, and the HIP
// kernels that evaluate them per matrix entry (fills) or contract a weight
// matrix against their parameter derivatives.
//
// Leaf formulas (the reference's, in log parameters):
//   Constant  k = c                           dk/dlog c = c        (src/kernel/constant.rs:12-16)
//   RBF       k = exp(-s / (2 l^2))           dk/dlog l = k s / l^2 (src/kernel/rbf.rs:18-20,254-256)
//   ARD-RBF   k = exp(-1/2 sum dx_d^2 / l_d^2) dk/dlog l_d = k dx_d^2 / l_d^2 (src/kernel/rbf_ard.rs:16,274-276)
//   Periodic  k = exp(-2 sin^2(pi r / p) / l^2), r = |x - x'|
//             dk/dlog l = k 4 sin^2 / l^2, dk/dlog p = k 4 sin cos a a / l^2, a = pi r / p
//             (src/kernel/periodic.rs:17-23,599-637)
// with s the squared Euclidean distance by direct differences
// (src/kernel/simd/dist.rs:7-19).
#include <algorithm>
#include <cstdio>
#include <cstring>

#include "fill.hpp"
#include "kernel_eval.hpp"

namespace gprx {

// ---------------------------------------------------------------------------
// Host: program parsing.
// ---------------------------------------------------------------------------
namespace {
struct Sop {
  std::vector<uint32_t> terms;  // bitmasks over leaves
};
}  // namespace

int parse_program(const gprx_kernel_program* prog, int d, HostProgram* out, std::string* why) {
  if (!prog || !prog->ops || prog->n_ops == 0) {
    *why = "empty kernel program";
    return GPRX_INVALID_HYPERPARAMETER;
  }
  HostProgram hp;
  hp.d = d;
  std::vector<Sop> stack;
  int poff = 0;
  for (size_t k = 0; k < prog->n_ops; ++k) {
    const int op = prog->ops[2 * k], arg = prog->ops[2 * k + 1];
    if (op >= GPRX_OP_CONST && op <= GPRX_OP_PERIODIC) {
      if (hp.nl >= kMaxLeaves) {
        *why = "kernel program has more than 8 leaves";
        return GPRX_UNSUPPORTED;
      }
      const int l = hp.nl++;
      hp.param_off[l] = poff;
      switch (op) {
        case GPRX_OP_CONST:
          hp.kind[l] = kLeafConst;
          poff += 1;
          break;
        case GPRX_OP_RBF:
          hp.kind[l] = kLeafRbf;
          poff += 1;
          break;
        case GPRX_OP_RBF_ARD:
          if (arg != d) {
            *why = "RBF-ARD lengthscale count differs from the input dimension";
            return GPRX_LENGTH_MISMATCH;
          }
          if (hp.ard_leaf >= 0) {
            *why = "more than one RBF-ARD leaf in a program";
            return GPRX_UNSUPPORTED;
          }
          hp.kind[l] = kLeafArd;
          hp.ard_leaf = l;
          poff += d;
          break;
        default:
          hp.kind[l] = kLeafPeriodic;
          poff += 2;
          break;
      }
      stack.push_back(Sop{{1u << l}});
    } else if (op == GPRX_OP_SUM || op == GPRX_OP_PROD) {
      if (arg < 1 || (size_t)arg > stack.size()) {
        *why = "malformed kernel program (operator arity)";
        return GPRX_INVALID_HYPERPARAMETER;
      }
      std::vector<Sop> kids(stack.end() - arg, stack.end());
      stack.resize(stack.size() - arg);
      Sop res;
      if (op == GPRX_OP_SUM) {
        for (auto& kd : kids) res.terms.insert(res.terms.end(), kd.terms.begin(), kd.terms.end());
      } else {
        res.terms = {0u};
        for (auto& kd : kids) {
          std::vector<uint32_t> nx;
          for (uint32_t a : res.terms)
            for (uint32_t b : kd.terms) nx.push_back(a | b);
          res.terms.swap(nx);
          if (res.terms.size() > (size_t)kMaxTerms) break;
        }
      }
      if (res.terms.size() > (size_t)kMaxTerms) {
        *why = "kernel program expands to more than 8 product terms";
        return GPRX_UNSUPPORTED;
      }
      stack.push_back(std::move(res));
    } else {
      *why = "unknown kernel program op";
      return GPRX_INVALID_HYPERPARAMETER;
    }
  }
  if (stack.size() != 1) {
    *why = "malformed kernel program (stack)";
    return GPRX_INVALID_HYPERPARAMETER;
  }
  hp.terms = stack[0].terms;
  hp.nt = (int)hp.terms.size();
  hp.n_params = poff;
  // Slots: the factors of each term, term after term.
  for (uint32_t mask : hp.terms) {
    bool first = true;
    for (int l = 0; l < hp.nl; ++l) {
      if (!((mask >> l) & 1u)) continue;
      if (hp.ns >= kMaxSlots) {
        *why = "kernel program expands to more than 8 leaf factors over its product terms";
        return GPRX_UNSUPPORTED;
      }
      if (first) hp.start |= 1u << hp.ns;
      first = false;
      hp.slot_leaf[hp.ns++] = l;
    }
    hp.end |= 1u << (hp.ns - 1);
  }
  *out = hp;
  return GPRX_OK;
}

void make_dev_program(const HostProgram& hp, const double* th, double diag_noise, DevProgram* P) {
  std::memset(P, 0, sizeof(*P));
  P->ns = hp.ns;
  P->d = hp.d;
  P->start = hp.start;
  P->end = hp.end;
  if (hp.ard_leaf >= 0) {
    const double* q = th + hp.param_off[hp.ard_leaf];
    for (int k = 0; k < hp.d; ++k) P->ard_w[k] = std::exp(-2.0 * q[k]);
  }
  for (int s = 0; s < hp.ns; ++s) {
    const int l = hp.slot_leaf[s];
    P->kind[s] = hp.kind[l];
    const double* q = th + hp.param_off[l];
    switch (hp.kind[l]) {
      case kLeafConst:
        P->a[s] = std::exp(q[0]);
        break;
      case kLeafRbf:
        P->a[s] = std::exp(-2.0 * q[0]);
        break;
      case kLeafArd:
        P->ard |= 1u << s;
        break;
      default:
        P->a[s] = M_PI * std::exp(-q[1]);
        P->b[s] = std::exp(-2.0 * q[0]);
        break;
    }
  }
  P->noise = diag_noise;
  for (int k = 0; k < kMaxSlots; ++k) {
    P->af[k] = (float)P->a[k];
    P->bf[k] = (float)P->b[k];
  }
  for (int k = 0; k < kMaxDim; ++k) P->ard_wf[k] = (float)P->ard_w[k];
}

void acc_to_grad(const HostProgram& hp, const double* th, const double* acc, double* grad) {
  for (int k = 0; k < hp.n_params; ++k) grad[k] = 0.0;
  for (int s = 0; s < hp.ns; ++s) {
    const int l = hp.slot_leaf[s];
    const int o = hp.param_off[l];
    const double a0 = acc[kAccPerLeaf * s], a1 = acc[kAccPerLeaf * s + 1];
    switch (hp.kind[l]) {
      case kLeafConst:
        grad[o] += a0;
        break;
      case kLeafRbf:
        grad[o] += a0 * std::exp(-2.0 * th[o]);
        break;
      case kLeafArd:
        break;  // summed over its slots on the device
      default: {
        const double b4 = 4.0 * std::exp(-2.0 * th[o]);
        grad[o] += a0 * b4;
        grad[o + 1] += a1 * b4;
        break;
      }
    }
  }
  if (hp.ard_leaf >= 0) {
    const int o = hp.param_off[hp.ard_leaf];
    for (int k = 0; k < hp.d; ++k)
      grad[o + k] = acc[kMaxSlots * kAccPerLeaf + k] * std::exp(-2.0 * th[o + k]);
  }
}

void set_error(gprx_error* err, int status, const std::string& msg, int stage, double jitter,
               size_t matrix_size) {
  if (!err) return;
  err->status = status;
  err->stage = stage;
  err->jitter = jitter;
  err->matrix_size = matrix_size;
  std::snprintf(err->message, sizeof(err->message), "%s", msg.c_str());
}

template <int DP, typename T, bool DUAL>
__global__ __launch_bounds__(256) void k_fill_lower(const DevProgram* __restrict__ Pg,
                                                    const double* __restrict__ x, long ldx, int n,
                                                    T* __restrict__ K, long ldk,
                                                    double* __restrict__ K64, int* info) {
  const DevProgram& P = *Pg;
  __shared__ double xr[DP][kTile];
  __shared__ double xc[DP][kTile];
  int ti, tj;
  tri_index_f(blockIdx.x, ti, tj);
  const int r0 = ti * kTile, c0 = tj * kTile;
  load_points<DP>(x, ldx, r0, xr);
  load_points<DP>(x, ldx, c0, xc);
  __syncthreads();
  const int i = threadIdx.x & 127, jh = threadIdx.x >> 7;
  const int gi = r0 + i;
  double xi[DP];
#pragma unroll
  for (int q = 0; q < DP; ++q) xi[q] = xr[q][i];
  bool bad = false;
  for (int jj = 0; jj < 64; ++jj) {
    const int j = jh + 2 * jj;
    const int gj = c0 + j;
    if (gi < gj) continue;
    double val;
    if (gi >= n || gj >= n) {
      val = gi == gj ? 1.0 : 0.0;
    } else {
      double dx[DP];
#pragma unroll
      for (int q = 0; q < DP; ++q) dx[q] = xi[q] - xc[q][j];
      Leaves L;
      val = eval_entry<DP, false>(P, dx, L);
      if (!isfinite(val)) bad = true;
      if (gi == gj) val += P.noise;
    }
    K[gi + (long)gj * ldk] = (T)val;
    if (DUAL) K64[gi + (long)gj * ldk] = val;
  }
  if (bad) atomicOr(info, 2);
}

template <int DP, typename T>
__global__ __launch_bounds__(256) void k_fill_cross(const DevProgram* __restrict__ Pg,
                                                    const double* __restrict__ a, long lda, int na,
                                                    const double* __restrict__ b, long ldb, int nb,
                                                    T* __restrict__ out, long ldo, int mt) {
  using R = T;  // f32 matrices: entries evaluated in f32
  const DevProgram& P = *Pg;
  __shared__ double xr[DP][kTile];
  __shared__ double xc[DP][kTile];
  const int ti = blockIdx.x % mt, tj = blockIdx.x / mt;
  const int r0 = ti * kTile, c0 = tj * kTile;
  load_points<DP>(a, lda, r0, xr);
  load_points<DP>(b, ldb, c0, xc);
  __syncthreads();
  const int i = threadIdx.x & 127, jh = threadIdx.x >> 7;
  const int gi = r0 + i;
  double xi[DP];
#pragma unroll
  for (int q = 0; q < DP; ++q) xi[q] = xr[q][i];
  for (int jj = 0; jj < 64; ++jj) {
    const int j = jh + 2 * jj;
    const int gj = c0 + j;
    double val = 0.0;
    if (gi < na && gj < nb) {
      R dx[DP];
#pragma unroll
      for (int q = 0; q < DP; ++q) dx[q] = (R)(xi[q] - xc[q][j]);
      LeavesT<R> L;
      val = eval_entry<DP, false>(P, dx, L);
    }
    out[gi + (long)gj * ldo] = (T)val;
  }
}

template <int DP>
__global__ void k_fill_diag(const DevProgram* __restrict__ Pg, int na, double* __restrict__ out) {
  const DevProgram& P = *Pg;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= na) return;
  double dx[DP];
#pragma unroll
  for (int q = 0; q < DP; ++q) dx[q] = 0.0;
  Leaves L;
  out[i] = eval_entry<DP, false>(P, dx, L);
}

template <int DP, typename T, int NS>
__global__ __launch_bounds__(256, contract_occupancy(DP, NS)) void k_contract_exact(const DevProgram* __restrict__ Pg,
                                                        const double* __restrict__ x, long ldx,
                                                        int n, const T* __restrict__ Ainv, long lda,
                                                        const double* __restrict__ alpha,
                                                        double* __restrict__ partial) {
  using R = T;  // f32 matrices: entries evaluated in f32
  const DevProgram& P = *Pg;
  __shared__ double xr[DP][kTile];
  __shared__ double xc[DP][kTile];
  __shared__ double ar[kTile], ac[kTile];
  __shared__ double red[4 * kMaxAcc];
  int ti, tj;
  tri_index_f(blockIdx.x, ti, tj);
  const int r0 = ti * kTile, c0 = tj * kTile;
  load_points<DP>(x, ldx, r0, xr);
  load_points<DP>(x, ldx, c0, xc);
  if (threadIdx.x < kTile) {
    ar[threadIdx.x] = alpha[r0 + threadIdx.x];
    ac[threadIdx.x] = alpha[c0 + threadIdx.x];
  }
  __syncthreads();
  const int i = threadIdx.x & 127, jh = threadIdx.x >> 7;
  const int gi = r0 + i;
  double xi[DP];
#pragma unroll
  for (int q = 0; q < DP; ++q) xi[q] = xr[q][i];
  const double ai = ar[i];
  // per-thread sums in the matrices' type (f32 sums of a thread's at most
  // 64 f32 entries, widened to f64 at the block reduction)
  using A = T;
  A acc[NS][kAccPerLeaf];
  A accd[DP];
#pragma unroll
  for (int l = 0; l < NS; ++l) acc[l][0] = acc[l][1] = 0;
#pragma unroll
  for (int q = 0; q < DP; ++q) accd[q] = 0;
  if (gi < n) {
    for (int jj = 0; jj < 64; ++jj) {
      const int j = jh + 2 * jj;
      const int gj = c0 + j;
      if (gj > gi || gj >= n) continue;
      double w = ai * ac[j] - (double)Ainv[gi + (long)gj * lda];
      if (gi != gj) w *= 2.0;
      R dx[DP];
#pragma unroll
      for (int q = 0; q < DP; ++q) dx[q] = (R)(xi[q] - xc[q][j]);
      LeavesT<R, NS> L;
      eval_entry<DP, true>(P, dx, L);
      accumulate_entry<DP>(P, dx, L, (A)w, acc, accd);
    }
  }
  store_partials<DP>(acc, accd, red, partial + (long)blockIdx.x * kMaxAcc);
}

template <int DP, typename T, int NS>
__global__ __launch_bounds__(256, contract_occupancy(DP, NS)) void k_contract_cross(const DevProgram* __restrict__ Pg,
                                                        const double* __restrict__ a, long lda,
                                                        int na, const double* __restrict__ b,
                                                        long ldb, int nb, const T* __restrict__ W,
                                                        long ldw, int mt,
                                                        double* __restrict__ partial,
                                                        const double* __restrict__ cv,
                                                        const double* __restrict__ rv, double crs) {
  using R = T;  // f32 matrices: entries evaluated in f32
  const DevProgram& P = *Pg;
  __shared__ double xr[DP][kTile];
  __shared__ double xc[DP][kTile];
  __shared__ double red[4 * kMaxAcc];
  const int ti = blockIdx.x % mt, tj = blockIdx.x / mt;
  const int r0 = ti * kTile, c0 = tj * kTile;
  load_points<DP>(a, lda, r0, xr);
  load_points<DP>(b, ldb, c0, xc);
  __syncthreads();
  const int i = threadIdx.x & 127, jh = threadIdx.x >> 7;
  const int gi = r0 + i;
  double xi[DP];
#pragma unroll
  for (int q = 0; q < DP; ++q) xi[q] = xr[q][i];
  // per-thread sums in the matrices' type (f32 sums of a thread's at most
  // 64 f32 entries, widened to f64 at the block reduction)
  using A = T;
  A acc[NS][kAccPerLeaf];
  A accd[DP];
#pragma unroll
  for (int l = 0; l < NS; ++l) acc[l][0] = acc[l][1] = 0;
#pragma unroll
  for (int q = 0; q < DP; ++q) accd[q] = 0;
  if (gi < na) {
    const double ci = cv ? crs * cv[gi] : 0.0;
    for (int jj = 0; jj < 64; ++jj) {
      const int j = jh + 2 * jj;
      const int gj = c0 + j;
      if (gj >= nb) continue;
      double w = (double)W[gi + (long)gj * ldw];
      if (cv) w -= ci * rv[gj];
      R dx[DP];
#pragma unroll
      for (int q = 0; q < DP; ++q) dx[q] = (R)(xi[q] - xc[q][j]);
      LeavesT<R, NS> L;
      eval_entry<DP, true>(P, dx, L);
      accumulate_entry<DP>(P, dx, L, (A)w, acc, accd);
    }
  }
  store_partials<DP>(acc, accd, red, partial + (long)blockIdx.x * kMaxAcc);
}

template <int DP>
__global__ __launch_bounds__(256) void k_contract_diag(const DevProgram* __restrict__ Pg, int na,
                                                       double w, double* __restrict__ partial) {
  const DevProgram& P = *Pg;
  __shared__ double red[4 * kMaxAcc];
  double acc[kMaxSlots][kAccPerLeaf];
  double accd[DP];
#pragma unroll
  for (int l = 0; l < kMaxSlots; ++l) acc[l][0] = acc[l][1] = 0.0;
#pragma unroll
  for (int q = 0; q < DP; ++q) accd[q] = 0.0;
  // k(x, x) does not depend on x for these stationary leaves: one entry
  // times the count.
  if (threadIdx.x == 0) {
    double dx[DP];
#pragma unroll
    for (int q = 0; q < DP; ++q) dx[q] = 0.0;
    Leaves L;
    eval_entry<DP, true>(P, dx, L);
    accumulate_entry<DP>(P, dx, L, w * (double)na, acc, accd);
  }
  store_partials<DP>(acc, accd, red, partial);
}

__global__ __launch_bounds__(256) void k_reduce_partials(const double* __restrict__ partial,
                                                         int tiles, double* __restrict__ out) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  double acc = 0.0;
  for (int t = threadIdx.x; t < tiles; t += 256) acc += partial[(long)t * kMaxAcc + k];
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0) out[k] = acc;
}

// ---------------------------------------------------------------------------
// Host launchers: dispatch on the padded dimension.
// ---------------------------------------------------------------------------
// The slot count a kernel is compiled for: the program's ns rounded up to
// 2, 4 or 8.
#define GPRX_DISPATCH_NS(ns, CALL)                      \
  if ((ns) <= 2) {                                      \
    constexpr int NS = 2;                               \
    CALL;                                               \
  } else if ((ns) <= 4) {                               \
    constexpr int NS = 4;                               \
    CALL;                                               \
  } else {                                              \
    constexpr int NS = kMaxSlots;                       \
    CALL;                                               \
  }

#define GPRX_DISPATCH_DIM(dp, CALL) \
  switch (dp) {                     \
    case 1: {                       \
      constexpr int DP = 1;         \
      CALL;                         \
    } break;                        \
    case 2: {                       \
      constexpr int DP = 2;         \
      CALL;                         \
    } break;                        \
    case 4: {                       \
      constexpr int DP = 4;         \
      CALL;                         \
    } break;                        \
    case 8: {                       \
      constexpr int DP = 8;         \
      CALL;                         \
    } break;                        \
    case 16: {                      \
      constexpr int DP = 16;        \
      CALL;                         \
    } break;                        \
    case 32: {                      \
      constexpr int DP = 32;        \
      CALL;                         \
    } break;                        \
    default:                        \
      throw DeviceError("unsupported padded dimension"); \
  }

template <typename T>
void fill_lower(int dp, const DevProgram* prog, DevPoints x, T* K, long ldk, double* K64, int* info,
                hipStream_t stream) {
  const int tiles = lower_tiles(x.ld);
  if (K64) {
    GPRX_DISPATCH_DIM(dp, hipLaunchKernelGGL((k_fill_lower<DP, T, true>), dim3(tiles), dim3(256),
                                             0, stream, prog, x.p, x.ld, x.rows, K, ldk, K64, info));
  } else {
    GPRX_DISPATCH_DIM(dp, hipLaunchKernelGGL((k_fill_lower<DP, T, false>), dim3(tiles), dim3(256),
                                             0, stream, prog, x.p, x.ld, x.rows, K, ldk, K64, info));
  }
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T>
void fill_cross(int dp, const DevProgram* prog, DevPoints a, DevPoints b, T* out, long ldo,
                hipStream_t stream) {
  const int mt = (int)(a.ld / kTile);
  const int tiles = cross_tiles(a.ld, b.ld);
  GPRX_DISPATCH_DIM(dp, hipLaunchKernelGGL((k_fill_cross<DP, T>), dim3(tiles), dim3(256), 0,
                                           stream, prog, a.p, a.ld, a.rows, b.p, b.ld, b.rows, out,
                                           ldo, mt));
  GPRX_HIP_CHECK(hipGetLastError());
}

void fill_diag(int dp, const DevProgram* prog, DevPoints a, double* out, hipStream_t stream) {
  GPRX_DISPATCH_DIM(dp, hipLaunchKernelGGL((k_fill_diag<DP>), dim3((a.rows + 255) / 256),
                                           dim3(256), 0, stream, prog, a.rows, out));
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T>
void contract_exact(int dp, int ns, const DevProgram* prog, DevPoints x, const T* Ainv, long lda,
                    const double* alpha, double* partial, hipStream_t stream) {
  const int tiles = lower_tiles(x.ld);
  GPRX_DISPATCH_DIM(dp, GPRX_DISPATCH_NS(ns, hipLaunchKernelGGL((k_contract_exact<DP, T, NS>),
                                                                dim3(tiles), dim3(256), 0, stream,
                                                                prog, x.p, x.ld, x.rows, Ainv, lda,
                                                                alpha, partial)));
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T>
void contract_cross(int dp, int ns, const DevProgram* prog, DevPoints a, DevPoints b, const T* W,
                    long ldw, double* partial, hipStream_t stream, const double* cv, const double* rv,
                    double crs) {
  const int mt = (int)(a.ld / kTile);
  const int tiles = cross_tiles(a.ld, b.ld);
  GPRX_DISPATCH_DIM(dp, GPRX_DISPATCH_NS(ns, hipLaunchKernelGGL((k_contract_cross<DP, T, NS>),
                                                                dim3(tiles), dim3(256), 0, stream,
                                                                prog, a.p, a.ld, a.rows, b.p, b.ld,
                                                                b.rows, W, ldw, mt, partial, cv, rv,
                                                                crs)));
  GPRX_HIP_CHECK(hipGetLastError());
}

void contract_diag(int dp, const DevProgram* prog, DevPoints a, double w, double* partial,
                   hipStream_t stream) {
  GPRX_DISPATCH_DIM(dp, hipLaunchKernelGGL((k_contract_diag<DP>), dim3(1), dim3(256), 0, stream,
                                           prog, a.rows, w, partial));
  GPRX_HIP_CHECK(hipGetLastError());
}

void reduce_partials(const double* partial, int tiles, double* out, hipStream_t stream) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(kMaxAcc), dim3(256), 0, stream, partial, tiles, out);
  GPRX_HIP_CHECK(hipGetLastError());
}

template void fill_lower<double>(int, const DevProgram*, DevPoints, double*, long, double*, int*,
                                 hipStream_t);
template void fill_lower<float>(int, const DevProgram*, DevPoints, float*, long, double*, int*,
                                hipStream_t);
template void fill_cross<double>(int, const DevProgram*, DevPoints, DevPoints, double*, long,
                                 hipStream_t);
template void fill_cross<float>(int, const DevProgram*, DevPoints, DevPoints, float*, long,
                                hipStream_t);
template void contract_exact<double>(int, int, const DevProgram*, DevPoints, const double*, long,
                                     const double*, double*, hipStream_t);
template void contract_exact<float>(int, int, const DevProgram*, DevPoints, const float*, long,
                                    const double*, double*, hipStream_t);
template void contract_cross<double>(int, int, const DevProgram*, DevPoints, DevPoints, const double*,
                                     long, double*, hipStream_t, const double*, const double*,
                                     double);
template void contract_cross<float>(int, int, const DevProgram*, DevPoints, DevPoints, const float*,
                                    long, double*, hipStream_t, const double*, const double*,
                                    double);

}  // namespace gprx
