// Dense linear-algebra kernels for gfx950 (CDNA4).
//
// Everything here works on 128 x 128 tiles of column-major matrices whose
// orders are padded to multiples of 128, so the hot loops carry no bounds
// checks.  Products run on the matrix cores through v_mfma_f64_16x16x4_f64
// (f64) and v_mfma_f32_16x16x4_f32 (f32); both share one lane layout:
//   srcA (16x4):  lane l holds X[l % 16][l / 16]
//   srcB (4x16):  lane l holds Y[l / 16][l % 16]
//   D (16x16):    lane l holds D[r(l, q)][l % 16], q = 0..3, with
//                 r = 4 * (l / 16) + q for f32 and r = l / 16 + 4 * q for f64
//                 (measured on gfx950 with an in-bounds probe of both builtins)
// The tile kernel feeds op(B) as srcA and op(A) as srcB, so D holds C^T and
// a lane's four results sit in four columns of C at one row: consecutive
// lanes store consecutive rows, which is the contiguous direction of a
// column-major C.
#include "dense.hpp"
#include "tile_gemm.hpp"

namespace gprx {

template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(256, kTileOccupancy) void k_gemm(GemmArgs<T> g) {
  using AV = typename Acc<T>::V;
  __shared__ TileSmem<T> sm;
  const int z = blockIdx.y;
  const T* __restrict__ A = g.A + (long)z * g.strideA;
  const T* __restrict__ B = g.B + (long)z * g.strideB;
  T* __restrict__ C = g.C + (long)z * g.strideC;
  const TileCoord tc = tile_coord(g.M, g.N, g.K, g.lower, g.order, g.krange);
  const int ti = tc.ti, tj = tc.tj;
  const int nk = tc.ke > tc.kb ? (tc.ke - tc.kb) / Cfg<T>::BK : 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  AV acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = AV{};
  const T* Ablk = TA ? A + (long)ti * kTile * g.lda : A + (long)ti * kTile;
  const T* Bblk = TB ? B + (long)tj * kTile : B + (long)tj * kTile * g.ldb;
  tile_mainloop<T, TA, TB>(Ablk, g.lda, Bblk, g.ldb, tc.kb, nk, sm, acc);

  const bool diag_tile = g.lower && ti == tj;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = wm * 64 + i * 16 + (lane & 15);
        const int c = wn * 64 + j * 16 + drow<T>(lane, q);
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
__device__ __forceinline__ void potrf_block_lds(const T* __restrict__ A, long lda, T* __restrict__ X, long ldx,
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
        // Every register index is a compile-time constant: a lane-dependent
        // index (d[r]) would move the arrays to scratch memory.
        double drr = 0.0, xrr = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          if (k <= r) s[(c0 + k) * DS + c0 + r] = d[k];
          if (k < r) s[(c0 + r) * DS + c0 + k] = xr[k];
          if (k == r) {
            drr = d[k];
            xrr = xr[k];
          }
        }
        xd[c0 + r] = xrr;
        logdiag[c0 + r] = log(drr);
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
      for (int q = 0; q < 4; ++q) s[(c0 + (lane & 15)) * DS + bi * 16 + drow<double>(lane, q)] = acc[q];
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
        const int row = bi * 16 + drow<double>(lane, q);
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
    for (int q = 0; q < 4; ++q) s[(kb * 16 + (lane & 15)) * DS + bi * 16 + drow<double>(lane, q)] = acc[q];
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
        const int r = bi * 16 + drow<double>(lane, q);
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

void prepare_dense_kernels() {
  set_diag_lds_once<double>();
  set_diag_lds_once<float>();
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
