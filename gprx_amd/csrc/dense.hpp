// Dense linear-algebra kernels of the gfx950 backend (declarations).
#pragma once

#include "common.hpp"

namespace gprx {

// K-range of a C tile (ti, tj), intersected: a triangular operand restricts
// the reduction to the blocks where it is non-zero.
enum : int {
  kKloRow = 1,  // k >= ti * 128
  kKloCol = 2,  // k >= tj * 128
  kKhiRow = 4,  // k <  (ti + 1) * 128
  kKhiCol = 8   // k <  (tj + 1) * 128
};

// Tile enumeration of a launch: which index runs fastest and whether the
// order is reversed, so the tiles with the longest reduction start first.
enum : int { kOrderRowFast = 0, kOrderRowFastRev = 1, kOrderColFast = 2, kOrderColFastRev = 3 };

template <typename T>
struct GemmArgs {
  const T* A;
  long lda;
  const T* B;
  long ldb;
  T* C;
  long ldc;
  int M, N, K;  // M, N multiples of 128; K multiple of 128
  double alpha, beta;
  int krange;  // kKlo*/kKhi* flags
  int lower;   // C square; only tiles ti >= tj; diagonal tiles keep r >= c
  int order;   // kOrder*
  int batch;   // independent problems along blockIdx.y (>= 1)
  long strideA, strideB, strideC;
  // Filled by gemm(): blockIdx.x -> (ti << 16 | tj), kNoTile for padding.
  const unsigned* sched;
};

constexpr unsigned kNoTile = 0xffffffffu;

// C = beta * C + alpha * op(A) op(B) over each tile's K-range.
// ta: A is stored K x M (op = transpose); tb: B is stored N x K.
template <typename T>
void gemm(bool ta, bool tb, const GemmArgs<T>& g, hipStream_t stream);

// The same product without the cached XCD schedule (whose first use of a
// shape allocates its device copy): each workgroup maps its own tile.
template <typename T>
void gemm_unscheduled(bool ta, bool tb, const GemmArgs<T>& g, hipStream_t stream);

// Factors the 128 x 128 diagonal block at A (lower part read; storage T,
// arithmetic f64) and writes X = L^-1 (lower, zero above) at X, the log of
// each pivot at logdiag[0..128).  A non-positive or non-finite pivot sets
// *info.  batch > 1 factors `batch` blocks strided along the diagonal.
template <typename T>
void potrf_diag(const T* A, long lda, T* X, long ldx, double* logdiag, int* info, int batch,
                long stride_a, long stride_x, long stride_log, hipStream_t stream);

// Diagonal factor over a list of blocks (batched problems): the block b sits
// at a_off[b] / x_off[b] (element offsets), with leading dimension ld[b];
// logdiag at log_off[b].
template <typename T>
void potrf_diag_list(T* A, T* X, const long* a_off, const long* x_off, const long* ld,
                     const long* log_off, double* logdiag, int* info, const int* info_idx, int count,
                     hipStream_t stream);

// out = M v for lower-triangular (tri = 1, reads r >= c) or full M (n x n or
// rows x cols), accumulated in f64.  abs_mode sums |M| v instead.  Uses
// part (cols/256 * rows doubles) as split-K scratch.
template <typename T>
void gemv_n(const T* M, long ld, int rows, int cols, const double* v, double* out, double* part,
            int tri, int strict, int abs_mode, hipStream_t stream);

// out_j = sum_i M_ij v_i over rows i >= j (tri = 1; strict = 1: i > j) or all
// rows; one wavefront per column.  accumulate: out += instead of =.
template <typename T>
void gemv_t(const T* M, long ld, int rows, int cols, const double* v, double* out, int tri,
            int strict, int abs_mode, int accumulate, hipStream_t stream);

// Products over the lower 128-tiles of an nt x nt tile matrix M (ld), one
// workgroup per tile reading it once, then a fixed-order sum of the tile
// partials (part: lower_tiles(nt) * 4 * 128 doubles):
//   kTmvRow   out = L v     (L: the stored lower tiles, diagonal tiles c <= r)
//   kTmvCol   out = L^T v
//   kTmvSym   out = A v     (A symmetric, stored by its lower triangle)
//   kTmvSymAbs  also out2 = |A| 1, the absolute row sums
enum { kTmvRow = 1, kTmvCol = 2, kTmvSym = 3, kTmvSymAbs = 7 };
template <typename T>
void tile_mv_lower(const T* M, long ld, int nt, int mode, const double* v, double* out, double* out2,
                   double* part, hipStream_t stream);

// out_j = sum_i M_ij^2 (column sums of squares, f64 accumulation).
template <typename T>
void col_sumsq(const T* M, long ld, int rows, int cols, double* out, hipStream_t stream);

// Elementwise f64 -> T copy / zero of an n-vector and of matrices.
template <typename T>
void zero(T* p, size_t count, hipStream_t stream);

// Sets kernel attributes (large dynamic LDS) once, outside any graph capture.
void prepare_dense_kernels();

}  // namespace gprx
