// Kernel-matrix fills and <W, dK/dtheta> contractions (declarations).
#pragma once

#include "common.hpp"

namespace gprx {

// Points on the device: a padded column-major (rows_pad x dim_pad) f64 array,
// zero beyond the true rows and dimensions.
struct DevPoints {
  const double* p = nullptr;
  long ld = 0;   // rows_pad
  int rows = 0;  // true rows
};

// Lower tiles of the training Gram K(x, x) + noise * I into K (storage T, ld
// ldk, order rows_pad), with identity on the padded diagonal.  When K64 is
// non-null the f64 values go there too (the mixed form keeps them for
// refinement residuals).  Non-finite values set bit 2 of *info.
template <typename T>
void fill_lower(int dim_pad, const DevProgram* prog, DevPoints x, T* K, long ldk, double* K64,
                int* info, hipStream_t stream);

// Full tiles of K(a, b) (a.rows x b.rows, padded to a.ld x b.ld) into out.
template <typename T>
void fill_cross(int dim_pad, const DevProgram* prog, DevPoints a, DevPoints b, T* out, long ldo,
                hipStream_t stream);

// k(a_i, a_i) for every point (no noise).
void fill_diag(int dim_pad, const DevProgram* prog, DevPoints a, double* out, hipStream_t stream);

// Exact-GPR gradient contraction over lower tiles: with
// W_ij = alpha_i alpha_j - Ainv_ij (counted twice off the diagonal), each
// tile writes its kMaxAcc accumulators (leaf slots, then ARD dimensions) to
// partial[tile * kMaxAcc + k].
// ns: the program's slot count (kernels are built for 2, 4 or 8 slots).
template <typename T>
void contract_exact(int dim_pad, int ns, const DevProgram* prog, DevPoints x, const T* Ainv,
                    long lda, const double* alpha, double* partial, hipStream_t stream);

// Contraction of a dense weight matrix W (a.rows x b.rows, f64 or f32, ld
// ldw) against dK(a, b): partial per tile as above.  Tiles are full
// (a.ld / 128) x (b.ld / 128).  With cv, the weight is W_ij - crs cv_i rv_j.
template <typename T>
void contract_cross(int dim_pad, int ns, const DevProgram* prog, DevPoints a, DevPoints b,
                    const T* W, long ldw, double* partial, hipStream_t stream,
                    const double* cv = nullptr, const double* rv = nullptr, double crs = 0.0);

// Contraction of a constant weight w against dk(a_i, a_i): one partial row.
void contract_diag(int dim_pad, const DevProgram* prog, DevPoints a, double w, double* partial,
                   hipStream_t stream);

// out[k] = sum_t partial[t * kMaxAcc + k] for k < kMaxAcc (fixed order).
void reduce_partials(const double* partial, int tiles, double* out, hipStream_t stream);

// Number of tiles the launches above write.
inline int lower_tiles(long rows_pad) {
  const int t = (int)(rows_pad / kTile);
  return t * (t + 1) / 2;
}
inline int cross_tiles(long a_pad, long b_pad) { return (int)(a_pad / kTile) * (int)(b_pad / kTile); }

}  // namespace gprx
