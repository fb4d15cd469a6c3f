// Dense linear-algebra kernels for gfx950 (CDNA4).
//
// Everything here works on 128 x 128 tiles of column-major matrices whose
// orders are padded to multiples of 128, so the hot loops carry no bounds
// checks.  Products run on the matrix cores through v_mfma_f64_16x16x4_f64
// (f64) and v_mfma_f32_16x16x4_f32 (f32); both share one lane layout:
//   srcA (16x4):  lane l holds X[l % 16][l / 16]
//   srcB (4x16):  lane l holds Y[l / 16][l % 16]
//   D (16x16):    lane l holds D[4 * (l / 16) + q][l % 16], q = 0..3
// The tile kernel feeds op(B) as srcA and op(A) as srcB, so D holds C^T and
// a lane's four results sit in four columns of C at one row: consecutive
// lanes store consecutive rows, which is the contiguous direction of a
// column-major C.
#include "dense.hpp"

namespace gprx {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

template <typename T>
struct Acc;
template <>
struct Acc<double> {
  using V = d4;
};
template <>
struct Acc<float> {
  using V = f4;
};

template <typename T>
struct Vec;
template <>
struct Vec<double> {
  using V = d2;
};
template <>
struct Vec<float> {
  using V = f4;
};

__device__ __forceinline__ d4 mfma16(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma16(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Tile-kernel configuration: K step and LDS row stride (elements).  The
// stride is 32 dwords past a multiple of 64 banks, so the four 16-lane groups
// of a fragment read alternate bank halves.
template <typename T>
struct Cfg;
template <>
struct Cfg<double> {
  static constexpr int BK = 16;
  static constexpr int LS = 144;
};
template <>
struct Cfg<float> {
  static constexpr int BK = 32;
  static constexpr int LS = 160;
};

// Lower-triangle enumeration by rows: bid = ti (ti + 1) / 2 + tj, ti >= tj.
__device__ __forceinline__ void tri_index(int bid, int& ti, int& tj) {
  int t = (int)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
  while ((t + 1) * (t + 2) / 2 <= bid) ++t;
  while (t * (t + 1) / 2 > bid) --t;
  ti = t;
  tj = bid - t * (t + 1) / 2;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Give every XCD one contiguous run of logical tiles so neighbouring
// tiles, which share A row-panels and B column-panels, share an L2.
__device__ __forceinline__ int xcd_remap(int bid, int total) {
  constexpr int kXcd = 8;
  if (total < kXcd) return bid;
  const int q = total / kXcd, r = total % kXcd;
  const int x = bid % kXcd, k = bid / kXcd;
  return x * q + (x < r ? x : r) + k;
}

template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs<T> g) {
  constexpr int BK = Cfg<T>::BK;
  constexpr int LS = Cfg<T>::LS;
  constexpr int VEC = 16 / sizeof(T);
  using V = typename Vec<T>::V;
  using AV = typename Acc<T>::V;
  __shared__ __attribute__((aligned(16))) T sA[2][BK * LS];
  __shared__ __attribute__((aligned(16))) T sB[2][BK * LS];

  const int z = blockIdx.y;
  const T* __restrict__ A = g.A + (long)z * g.strideA;
  const T* __restrict__ B = g.B + (long)z * g.strideB;
  T* __restrict__ C = g.C + (long)z * g.strideC;

  const int mt = g.M / kTile, nt = g.N / kTile;
  const int total = g.lower ? mt * (mt + 1) / 2 : mt * nt;
  int bid = xcd_remap(blockIdx.x, total);
  int ti, tj;
  if (g.lower) {
    if (g.order & 1) bid = total - 1 - bid;
    tri_index(bid, ti, tj);
  } else {
    if (g.order & 1) bid = total - 1 - bid;
    if (g.order < 2) {
      ti = bid % mt;
      tj = bid / mt;
    } else {
      tj = bid % nt;
      ti = bid / nt;
    }
  }
  int kb = 0, ke = g.K;
  if (g.krange & kKloRow) kb = max(kb, ti * kTile);
  if (g.krange & kKloCol) kb = max(kb, tj * kTile);
  if (g.krange & kKhiRow) ke = min(ke, (ti + 1) * kTile);
  if (g.krange & kKhiCol) ke = min(ke, (tj + 1) * kTile);
  const int nk = ke > kb ? (ke - kb) / BK : 0;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  AV acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = AV{};

  // Global tile origins.
  const T* Ablk = TA ? A + (long)ti * kTile * g.lda : A + (long)ti * kTile;
  const T* Bblk = TB ? B + (long)tj * kTile : B + (long)tj * kTile * g.ldb;
  V ra[4], rb[4];

  // A "row-contiguous" operand has element (mn, k) at P[mn + k * ld]; a
  // "k-contiguous" one at P[k + mn * ld].
  auto gload = [&](int k0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int e = tid + 256 * v;
      if (!TA) {
        const int k = e / (kTile / VEC), mn = (e % (kTile / VEC)) * VEC;
        ra[v] = *reinterpret_cast<const V*>(Ablk + mn + (long)(k0 + k) * g.lda);
      } else {
        const int mn = e / (BK / VEC), k = (e % (BK / VEC)) * VEC;
        ra[v] = *reinterpret_cast<const V*>(Ablk + (k0 + k) + (long)mn * g.lda);
      }
      if (TB) {
        const int k = e / (kTile / VEC), mn = (e % (kTile / VEC)) * VEC;
        rb[v] = *reinterpret_cast<const V*>(Bblk + mn + (long)(k0 + k) * g.ldb);
      } else {
        const int mn = e / (BK / VEC), k = (e % (BK / VEC)) * VEC;
        rb[v] = *reinterpret_cast<const V*>(Bblk + (k0 + k) + (long)mn * g.ldb);
      }
    }
  };
  auto lstore = [&](int buf) {
    T* a = sA[buf];
    T* b = sB[buf];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int e = tid + 256 * v;
      if (!TA) {
        const int k = e / (kTile / VEC), mn = (e % (kTile / VEC)) * VEC;
        *reinterpret_cast<V*>(a + k * LS + mn) = ra[v];
      } else {
        const int mn = e / (BK / VEC), k = (e % (BK / VEC)) * VEC;
#pragma unroll
        for (int c = 0; c < VEC; ++c) a[(k + c) * LS + mn] = ra[v][c];
      }
      if (TB) {
        const int k = e / (kTile / VEC), mn = (e % (kTile / VEC)) * VEC;
        *reinterpret_cast<V*>(b + k * LS + mn) = rb[v];
      } else {
        const int mn = e / (BK / VEC), k = (e % (BK / VEC)) * VEC;
#pragma unroll
        for (int c = 0; c < VEC; ++c) b[(k + c) * LS + mn] = rb[v][c];
      }
    }
  };

  if (nk > 0) {
    gload(kb);
    lstore(0);
  }
  __syncthreads();
  for (int it = 0; it < nk; ++it) {
    const int cur = it & 1;
    if (it + 1 < nk) gload(kb + (it + 1) * BK);
    const T* a = sA[cur];
    const T* b = sB[cur];
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      T af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = a[kr * LS + wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = b[kr * LS + wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mfma16(bf[j], af[i], acc[i][j]);
    }
    if (it + 1 < nk) lstore(cur ^ 1);
    __syncthreads();
  }

  const bool diag_tile = g.lower && ti == tj;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = wm * 64 + i * 16 + (lane & 15);
        const int c = wn * 64 + j * 16 + 4 * (lane >> 4) + q;
        if (diag_tile && r < c) continue;
        T* p = C + (long)(tj * kTile + c) * g.ldc + ti * kTile + r;
        double v = g.alpha * (double)acc[i][j][q];
        if (g.beta != 0.0) v += g.beta * (double)(*p);
        *p = (T)v;
      }
}

template <typename T>
void gemm(bool ta, bool tb, const GemmArgs<T>& g, hipStream_t stream) {
  const int mt = g.M / kTile, nt = g.N / kTile;
  const int total = g.lower ? mt * (mt + 1) / 2 : mt * nt;
  if (total == 0) return;
  dim3 grid(total, g.batch > 0 ? g.batch : 1);
  if (!ta && !tb)
    hipLaunchKernelGGL((k_gemm<T, false, false>), grid, dim3(256), 0, stream, g);
  else if (!ta && tb)
    hipLaunchKernelGGL((k_gemm<T, false, true>), grid, dim3(256), 0, stream, g);
  else if (ta && !tb)
    hipLaunchKernelGGL((k_gemm<T, true, false>), grid, dim3(256), 0, stream, g);
  else
    hipLaunchKernelGGL((k_gemm<T, true, true>), grid, dim3(256), 0, stream, g);
  GPRX_HIP_CHECK(hipGetLastError());
}

template void gemm<double>(bool, bool, const GemmArgs<double>&, hipStream_t);
template void gemm<float>(bool, bool, const GemmArgs<float>&, hipStream_t);

// ---------------------------------------------------------------------------
// 128 x 128 diagonal block: Cholesky + inverse in LDS, f64 arithmetic.
//
// Layout in LDS (column stride DS): L lower (r >= c) at s[c * DS + r]; the
// inverse X = L^-1 stored transposed above the diagonal, X[r][c] (r > c) at
// s[r * DS + c]; X[r][r] in xd[r].  Steps, per 16-column panel:
//   (a) wave 0 factors the 16 x 16 diagonal sub-block in registers (lane r
//       owns row r; pivots and columns move by cross-lane shuffles) and
//       inverts it by forward substitution;
//   (b) panel rows below: L_i = A_i X_D^T on MFMA;
//   (c) trailing update A_ij -= L_i L_j^T (lower) on MFMA.
// Then Y_ik = X_D[i] L_ik in place, and X_ij = -sum_k Y_ik X_kj by
// diagonal distance, all on MFMA.
// ---------------------------------------------------------------------------
constexpr int DS = 129;

__device__ __forceinline__ double x_at(const double* s, const double* xd, int r, int c) {
  return r > c ? s[r * DS + c] : (r == c ? xd[r] : 0.0);
}

__device__ __forceinline__ void tri_index_small(int t, int& u, int& v) {
  u = 0;
  while ((u + 1) * (u + 2) / 2 <= t) ++u;
  v = t - u * (u + 1) / 2;
}

template <typename T>
__device__ void potrf_block_lds(const T* __restrict__ A, long lda, T* __restrict__ X, long ldx,
                                double* __restrict__ logdiag, int* info, double* s, double* xd,
                                int* sbad) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) *sbad = 0;
  for (int e = tid; e < kTile * kTile; e += 256) {
    const int r = e & 127, c = e >> 7;
    s[c * DS + r] = (r >= c) ? (double)A[r + (long)c * lda] : 0.0;
  }
  __syncthreads();

  for (int jb = 0; jb < 8; ++jb) {
    const int c0 = jb * 16;
    if (wave == 0) {
      const int r = lane & 15;
      double d[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) d[k] = (k <= r) ? s[(c0 + k) * DS + c0 + r] : 0.0;
      bool bad = false;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        double pj = __shfl(d[j], j);
        if (!(pj > 0.0) || !isfinite(pj)) {
          bad = true;
          pj = 1.0;
        }
        const double ljj = sqrt(pj);
        if (r == j)
          d[j] = ljj;
        else if (r > j)
          d[j] /= ljj;
#pragma unroll
        for (int c = j + 1; c < 16; ++c) {
          const double lcj = __shfl(d[j], c);
          if (r >= c) d[c] -= d[j] * lcj;
        }
      }
      if (bad && lane == 0) *sbad = 1;
      double xr[16], acc[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        xr[c] = 0.0;
        acc[c] = 0.0;
      }
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (r == k) {
          const double inv = 1.0 / d[k];
#pragma unroll
          for (int c = 0; c < k; ++c) xr[c] = -acc[c] * inv;
          xr[k] = inv;
        }
#pragma unroll
        for (int c = 0; c <= k; ++c) {
          const double xkc = __shfl(xr[c], k);
          if (r > k) acc[c] += d[k] * xkc;
        }
      }
      if (lane < 16) {
#pragma unroll
        for (int k = 0; k < 16; ++k)
          if (k <= r) s[(c0 + k) * DS + c0 + r] = d[k];
#pragma unroll
        for (int c = 0; c < 16; ++c)
          if (c < r) s[(c0 + r) * DS + c0 + c] = xr[c];
        xd[c0 + r] = xr[r];
        logdiag[c0 + r] = log(d[r]);
      }
    }
    __syncthreads();
    // (b) panel
    for (int bi = jb + 1 + wave; bi < 8; bi += 4) {
      d4 acc = d4{};
#pragma unroll
      for (int kk = 0; kk < 16; kk += 4) {
        const int kr = kk + (lane >> 4);
        const int jr = lane & 15;
        const double a = s[(c0 + kr) * DS + bi * 16 + jr];
        const double b = x_at(s, xd, c0 + jr, c0 + kr);
        acc = mfma16(a, b, acc);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) s[(c0 + (lane & 15)) * DS + bi * 16 + 4 * (lane >> 4) + q] = acc[q];
    }
    __syncthreads();
    // (c) trailing update
    const int nb = 7 - jb;
    const int count = nb * (nb + 1) / 2;
    for (int t = wave; t < count; t += 4) {
      int u, v;
      tri_index_small(t, u, v);
      const int bi = jb + 1 + u, bj = jb + 1 + v;
      d4 acc = d4{};
#pragma unroll
      for (int kk = 0; kk < 16; kk += 4) {
        const int kr = kk + (lane >> 4);
        const double a = s[(c0 + kr) * DS + bi * 16 + (lane & 15)];
        const double b = s[(c0 + kr) * DS + bj * 16 + (lane & 15)];
        acc = mfma16(a, b, acc);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = bi * 16 + 4 * (lane >> 4) + q;
        const int col = bj * 16 + (lane & 15);
        if (bi != bj || row >= col) s[col * DS + row] -= acc[q];
      }
    }
    __syncthreads();
  }

  // Y_{bi,kb} = X_D[bi] L_{bi,kb}, in place over L (28 strictly-lower blocks).
  for (int t = wave; t < 28; t += 4) {
    int u, v;
    tri_index_small(t, u, v);
    const int bi = u + 1, kb = v;
    d4 acc = d4{};
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      const int kr = kk + (lane >> 4);
      const double a = x_at(s, xd, bi * 16 + (lane & 15), bi * 16 + kr);
      const double b = s[(kb * 16 + (lane & 15)) * DS + bi * 16 + kr];
      acc = mfma16(a, b, acc);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) s[(kb * 16 + (lane & 15)) * DS + bi * 16 + 4 * (lane >> 4) + q] = acc[q];
  }
  __syncthreads();
  // X_{bi,bj} = -sum_{k=bj}^{bi-1} Y_{bi,k} X_{k,bj}, by distance bi - bj.
  for (int dist = 1; dist < 8; ++dist) {
    for (int bj = wave; bj < 8 - dist; bj += 4) {
      const int bi = bj + dist;
      d4 acc = d4{};
      for (int k = bj; k < bi; ++k) {
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
          const int kr = kk + (lane >> 4);
          const double a = s[(k * 16 + kr) * DS + bi * 16 + (lane & 15)];
          const double b = x_at(s, xd, k * 16 + kr, bj * 16 + (lane & 15));
          acc = mfma16(a, b, acc);
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = bi * 16 + 4 * (lane >> 4) + q;
        const int c = bj * 16 + (lane & 15);
        s[r * DS + c] = -acc[q];
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < kTile * kTile; e += 256) {
    const int r = e & 127, c = e >> 7;
    X[r + (long)c * ldx] = (T)x_at(s, xd, r, c);
  }
  if (tid == 0 && *sbad) atomicExch(info, 1);
}

template <typename T>
__global__ __launch_bounds__(256) void k_potrf_diag(const T* A, long lda, T* X, long ldx,
                                                    double* logdiag, int* info, long sa, long sx,
                                                    long slog) {
  extern __shared__ double smem[];
  double* s = smem;
  double* xd = smem + kTile * DS;
  int* sbad = reinterpret_cast<int*>(xd + kTile);
  const long z = blockIdx.x;
  potrf_block_lds<T>(A + z * sa, lda, X + z * sx, ldx, logdiag + z * slog, info, s, xd, sbad);
}

template <typename T>
__global__ __launch_bounds__(256) void k_potrf_diag_list(T* A, T* X, const long* a_off,
                                                         const long* x_off, const long* ld,
                                                         const long* log_off, double* logdiag,
                                                         int* info, const int* info_idx) {
  extern __shared__ double smem[];
  double* s = smem;
  double* xd = smem + kTile * DS;
  int* sbad = reinterpret_cast<int*>(xd + kTile);
  const int b = blockIdx.x;
  potrf_block_lds<T>(A + a_off[b], ld[b], X + x_off[b], ld[b], logdiag + log_off[b],
                     info + info_idx[b], s, xd, sbad);
}

constexpr size_t kDiagLds = (size_t)kTile * DS * sizeof(double) + kTile * sizeof(double) + 16;

template <typename T>
static void set_diag_lds_once() {
  static bool done = false;
  if (!done) {
    GPRX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_potrf_diag<T>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDiagLds));
    GPRX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_potrf_diag_list<T>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDiagLds));
    done = true;
  }
}

template <typename T>
void potrf_diag(const T* A, long lda, T* X, long ldx, double* logdiag, int* info, int batch,
                long stride_a, long stride_x, long stride_log, hipStream_t stream) {
  set_diag_lds_once<T>();
  hipLaunchKernelGGL((k_potrf_diag<T>), dim3(batch), dim3(256), kDiagLds, stream, A, lda, X, ldx,
                     logdiag, info, stride_a, stride_x, stride_log);
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T>
void potrf_diag_list(T* A, T* X, const long* a_off, const long* x_off, const long* ld,
                     const long* log_off, double* logdiag, int* info, const int* info_idx, int count,
                     hipStream_t stream) {
  if (count <= 0) return;
  set_diag_lds_once<T>();
  hipLaunchKernelGGL((k_potrf_diag_list<T>), dim3(count), dim3(256), kDiagLds, stream, A, X, a_off,
                     x_off, ld, log_off, logdiag, info, info_idx);
  GPRX_HIP_CHECK(hipGetLastError());
}

template void potrf_diag<double>(const double*, long, double*, long, double*, int*, int, long, long,
                                 long, hipStream_t);
template void potrf_diag<float>(const float*, long, float*, long, double*, int*, int, long, long,
                                long, hipStream_t);
template void potrf_diag_list<double>(double*, double*, const long*, const long*, const long*,
                                      const long*, double*, int*, const int*, int, hipStream_t);
template void potrf_diag_list<float>(float*, float*, const long*, const long*, const long*,
                                     const long*, double*, int*, const int*, int, hipStream_t);

// ---------------------------------------------------------------------------
// Matrix-vector products (HBM-bound; f64 accumulation, fixed order).
// ---------------------------------------------------------------------------
constexpr int kGemvChunk = 256;

template <typename T>
__global__ __launch_bounds__(256) void k_gemv_n_part(const T* __restrict__ M, long ld, int rows,
                                                     int cols, const double* __restrict__ v,
                                                     double* __restrict__ part, int tri, int strict,
                                                     int abs_mode) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * kGemvChunk;
  if (i >= rows) return;
  int c1 = min(cols, c0 + kGemvChunk);
  if (tri) c1 = min(c1, strict ? i : i + 1);
  double acc = 0.0;
  for (int c = c0; c < c1; ++c) {
    const double m = (double)M[i + (long)c * ld];
    acc += (abs_mode ? fabs(m) : m) * v[c];
  }
  part[(long)blockIdx.y * rows + i] = acc;
}

__global__ void k_sum_parts(const double* __restrict__ part, int rows, int nparts,
                            double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= rows) return;
  double acc = 0.0;
  for (int p = 0; p < nparts; ++p) acc += part[(long)p * rows + i];
  out[i] = acc;
}

template <typename T>
void gemv_n(const T* M, long ld, int rows, int cols, const double* v, double* out, double* part,
            int tri, int strict, int abs_mode, hipStream_t stream) {
  const int nparts = (cols + kGemvChunk - 1) / kGemvChunk;
  dim3 grid((rows + 255) / 256, nparts);
  hipLaunchKernelGGL((k_gemv_n_part<T>), grid, dim3(256), 0, stream, M, ld, rows, cols, v, part,
                     tri, strict, abs_mode);
  hipLaunchKernelGGL(k_sum_parts, dim3((rows + 255) / 256), dim3(256), 0, stream, part, rows, nparts,
                     out);
  GPRX_HIP_CHECK(hipGetLastError());
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gemv_t(const T* __restrict__ M, long ld, int rows,
                                                int cols, const double* __restrict__ v,
                                                double* __restrict__ out, int tri, int strict,
                                                int abs_mode, int accumulate, int sumsq) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= cols) return;
  const int r0 = tri ? (strict ? j + 1 : j) : 0;
  const T* col = M + (long)j * ld;
  double acc = 0.0;
  for (int i = r0 + lane; i < rows; i += 64) {
    const double m = (double)col[i];
    if (sumsq)
      acc += m * m;
    else
      acc += (abs_mode ? fabs(m) : m) * v[i];
  }
  acc = wave_sum(acc);
  if (lane == 0) out[j] = accumulate ? out[j] + acc : acc;
}

template <typename T>
void gemv_t(const T* M, long ld, int rows, int cols, const double* v, double* out, int tri,
            int strict, int abs_mode, int accumulate, hipStream_t stream) {
  hipLaunchKernelGGL((k_gemv_t<T>), dim3((cols + 3) / 4), dim3(256), 0, stream, M, ld, rows, cols,
                     v, out, tri, strict, abs_mode, accumulate, 0);
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T>
void col_sumsq(const T* M, long ld, int rows, int cols, double* out, hipStream_t stream) {
  hipLaunchKernelGGL((k_gemv_t<T>), dim3((cols + 3) / 4), dim3(256), 0, stream, M, ld, rows, cols,
                     (const double*)nullptr, out, 0, 0, 0, 0, 1);
  GPRX_HIP_CHECK(hipGetLastError());
}

template void gemv_n<double>(const double*, long, int, int, const double*, double*, double*, int,
                             int, int, hipStream_t);
template void gemv_n<float>(const float*, long, int, int, const double*, double*, double*, int, int,
                            int, hipStream_t);
template void gemv_t<double>(const double*, long, int, int, const double*, double*, int, int, int,
                             int, hipStream_t);
template void gemv_t<float>(const float*, long, int, int, const double*, double*, int, int, int,
                            int, hipStream_t);
template void col_sumsq<double>(const double*, long, int, int, double*, hipStream_t);
template void col_sumsq<float>(const float*, long, int, int, double*, hipStream_t);

template <typename T>
__global__ void k_zero(T* p, size_t count) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t k = i; k < count; k += stride) p[k] = T(0);
}

template <typename T>
void zero(T* p, size_t count, hipStream_t stream) {
  if (count == 0) return;
  size_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((k_zero<T>), dim3((unsigned)blocks), dim3(256), 0, stream, p, count);
  GPRX_HIP_CHECK(hipGetLastError());
}
template void zero<double>(double*, size_t, hipStream_t);
template void zero<float>(float*, size_t, hipStream_t);

__global__ void k_mfma_probe(const double* A, const double* B, double* C) {
  const int l = threadIdx.x;
  const double a = A[(l % 16) * 4 + l / 16];  // A row-major 16 x 4
  const double b = B[(l / 16) * 16 + l % 16];  // B row-major 4 x 16
  d4 acc = d4{};
  acc = mfma16(a, b, acc);
#pragma unroll
  for (int q = 0; q < 4; ++q) C[(4 * (l / 16) + q) * 16 + l % 16] = acc[q];  // C row-major
}

void mfma_probe(const double* A, const double* B, double* C, hipStream_t stream) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, stream, A, B, C);
  GPRX_HIP_CHECK(hipGetLastError());
}

}  // namespace gprx
