// Exact NLML Hessian pieces on the device (see hessian.hip).
#pragma once

#include "fill.hpp"

namespace gprx {

// Per slot of a program: the first theta index of its leaf and how many
// parameters that leaf owns.
struct HessSlots {
  int po[kMaxSlots];
  int cnt[kMaxSlots];
};
HessSlots hess_slots(const HostProgram& hp);

// D = dK/dtheta_param (kernel parameters only), both triangles, i, j < n.
void dfill(int dim_pad, const DevProgram* prog, const HessSlots& H, DevPoints x, int param,
           double* D, long ldd, hipStream_t s);
// partial[tile] (lower_tiles(x.ld) of them) = sum over the tile's lower
// entries of W d2K/dtheta_i dtheta_j, W = alpha alpha^T - Ainv, doubled off
// the diagonal.
void contract2(int dim_pad, const DevProgram* prog, const HessSlots& H, DevPoints x,
               const double* Ainv, long lda, const double* alpha, int pi, int pj, double* partial,
               hipStream_t s);
// partial[c] (n of them) = sum_{r >= c} w A_rc B_rc, w = 1 on the diagonal, 2 below.
void dot_lower(const double* A, const double* B, long ld, int n, double* partial, hipStream_t s);
// D_ii = v for i < n.
void set_diag(double* D, long ld, int n, double v, hipStream_t s);

}  // namespace gprx
