// Device side of the exact NLML Hessian (ExactFit::fill_hessian_from_factor,
// src/gpr/exact_fit.rs:484-596): for parameters i, j of the kernel program,
//   second order  -1/2 <W, d2A/dtheta_i dtheta_j>,  W = alpha alpha^T - A^-1,
//   first order   -1/2 tr(S_i S_j) + v_i . v_j,  S_i = L^-1 A_i L^-T,
//                 v_i = L^-1 A_i alpha,  A_i = dA/dtheta_i.
// Derivatives are with respect to the log-parameters.  A leaf value v with
// log-parameters t has dv/dt_a = v phi_a and d2v/dt_a dt_b = v (phi_a phi_b
// + c_ab); a product term T gets d2T = T (Phi_i Phi_j + c_ij), with c_ij
// nonzero only when one leaf owns both parameters:
//   constant   phi = 1                      c = 0
//   rbf        phi = s / l^2                c = -2 phi
//   ard        phi_d = dx_d^2 / l_d^2       c_dd = -2 phi_d
//   periodic   phi_l = 4 sin^2(al) / l^2,   phi_p = 4 sin(al) cos(al) al / l^2
//              c_ll = -2 phi_l,  c_lp = -2 phi_p,
//              c_pp = -phi_p - 4 al^2 cos(2 al) / l^2      (al = pi r / p)
// (the leaf derivatives of src/kernel/{constant,rbf,rbf_ard,periodic}.rs
// differentiated once more).
#include "hessian.hpp"
#include "kernel_eval.hpp"

namespace gprx {

namespace {

// dK/dtheta_i, dK/dtheta_j and d2K/dtheta_i dtheta_j at one entry (f64).
template <int DP>
__device__ __forceinline__ void entry_derivs(const DevProgram& P, const HessSlots& H,
                                             const double (&dx)[DP], int pi, int pj, double* d1i,
                                             double* d1j, double* d2) {
  double s = 0.0, r2 = 0.0;
#pragma unroll
  for (int q = 0; q < DP; ++q) {
    const double t = dx[q] * dx[q];
    s += t;
    r2 += t * P.ard_w[q];
  }
  double T = 1.0, Fi = 0.0, Fj = 0.0, C = 0.0;
  double gi = 0.0, gj = 0.0, h = 0.0;
#pragma unroll 1
  for (int sl = 0; sl < P.ns; ++sl) {
    if ((P.start >> sl) & 1u) {
      T = 1.0;
      Fi = Fj = C = 0.0;
    }
    const int kd = P.kind[sl];
    const int li = pi - H.po[sl], lj = pj - H.po[sl];
    const bool own_i = li >= 0 && li < H.cnt[sl], own_j = lj >= 0 && lj < H.cnt[sl];
    double v, phi_i = 0.0, phi_j = 0.0, c = 0.0;
    if (kd == kLeafConst) {
      v = P.a[sl];
      phi_i = 1.0;
      phi_j = 1.0;
    } else if (kd == kLeafRbf) {
      v = exp(-0.5 * s * P.a[sl]);
      phi_i = phi_j = s * P.a[sl];
      c = -2.0 * phi_i;
    } else if (kd == kLeafArd) {
      v = exp(-0.5 * r2);
#pragma unroll
      for (int q = 0; q < DP; ++q) {
        const double f = dx[q] * dx[q] * P.ard_w[q];
        phi_i = (q == li) ? f : phi_i;
        phi_j = (q == lj) ? f : phi_j;
      }
      c = (li == lj) ? -2.0 * phi_i : 0.0;
    } else {
      const double r = sqrt(s);
      const double al = r * P.a[sl], b = P.b[sl];
      double sn, cs;
      sincos(al, &sn, &cs);
      v = exp(-2.0 * sn * sn * b);
      const double fl = 4.0 * sn * sn * b, fp = 4.0 * b * sn * cs * al;
      phi_i = li == 0 ? fl : fp;
      phi_j = lj == 0 ? fl : fp;
      if (li == 0 && lj == 0)
        c = -2.0 * fl;
      else if (li == 1 && lj == 1)
        c = -fp - 4.0 * b * al * al * (cs * cs - sn * sn);
      else
        c = -2.0 * fp;
    }
    T *= v;
    if (own_i) Fi = phi_i;
    if (own_j) Fj = phi_j;
    if (own_i && own_j) C = c;
    if ((P.end >> sl) & 1u) {
      gi += T * Fi;
      gj += T * Fj;
      h += T * (Fi * Fj + C);
    }
  }
  *d1i = gi;
  *d1j = gj;
  *d2 = h;
}

// D = dK/dtheta_p over the lower tiles, written to both triangles (the pad
// stays as the caller cleared it).
template <int DP>
__global__ __launch_bounds__(256) void k_dfill(const DevProgram* __restrict__ Pg, HessSlots H,
                                               const double* __restrict__ x, long ldx, int n,
                                               int param, double* __restrict__ D, long ldd) {
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
  if (gi >= n) return;
  for (int jj = 0; jj < 64; ++jj) {
    const int j = jh + 2 * jj;
    const int gj = c0 + j;
    if (gj > gi || gj >= n) continue;
    double dx[DP];
#pragma unroll
    for (int q = 0; q < DP; ++q) dx[q] = xr[q][i] - xc[q][j];
    double d1, d1b, d2;
    entry_derivs<DP>(P, H, dx, param, param, &d1, &d1b, &d2);
    D[gi + (long)gj * ldd] = d1;
    D[gj + (long)gi * ldd] = d1;
  }
}

// partial[tile] = sum over the tile's lower entries of W_e d2K_e/dtheta_i
// dtheta_j, W = alpha alpha^T - Ainv counted twice off the diagonal.
template <int DP>
__global__ __launch_bounds__(256) void k_contract2(const DevProgram* __restrict__ Pg, HessSlots H,
                                                   const double* __restrict__ x, long ldx, int n,
                                                   const double* __restrict__ Ainv, long lda,
                                                   const double* __restrict__ alpha, int pi, int pj,
                                                   double* __restrict__ partial) {
  const DevProgram& P = *Pg;
  __shared__ double xr[DP][kTile];
  __shared__ double xc[DP][kTile];
  __shared__ double ar[kTile], ac[kTile];
  __shared__ double red[4];
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
  double acc = 0.0;
  if (gi < n) {
    for (int jj = 0; jj < 64; ++jj) {
      const int j = jh + 2 * jj;
      const int gj = c0 + j;
      if (gj > gi || gj >= n) continue;
      double w = ar[i] * ac[j] - Ainv[gi + (long)gj * lda];
      if (gi != gj) w *= 2.0;
      double dx[DP];
#pragma unroll
      for (int q = 0; q < DP; ++q) dx[q] = xr[q][i] - xc[q][j];
      double d1, d1b, d2;
      entry_derivs<DP>(P, H, dx, pi, pj, &d1, &d1b, &d2);
      acc += w * d2;
    }
  }
  const double v = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[wave_index()] = v;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// partial[c] = sum over rows r >= c of weight * A_rc B_rc (weight 1 on the
// diagonal, 2 below it), one block per column.
__global__ __launch_bounds__(256) void k_dot_lower(const double* __restrict__ A,
                                                   const double* __restrict__ B, long ld, int n,
                                                   double* __restrict__ partial) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  const long base = (long)c * ld;
  double acc = 0.0;
  for (int r = c + threadIdx.x; r < n; r += 256) {
    const double w = r == c ? 1.0 : 2.0;
    acc += w * A[base + r] * B[base + r];
  }
  const double v = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[wave_index()] = v;
  __syncthreads();
  if (threadIdx.x == 0) partial[c] = ((red[0] + red[1]) + red[2]) + red[3];
}

// D_ii = v for i < n (the rest of D untouched).
__global__ void k_set_diag(double* D, long ld, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) D[i + (long)i * ld] = v;
}

#define GPRX_HESS_DISPATCH(dp, CALL) \
  switch (dp) {                      \
    case 1: {                        \
      constexpr int DP = 1;          \
      CALL;                          \
    } break;                         \
    case 2: {                        \
      constexpr int DP = 2;          \
      CALL;                          \
    } break;                         \
    case 4: {                        \
      constexpr int DP = 4;          \
      CALL;                          \
    } break;                         \
    case 8: {                        \
      constexpr int DP = 8;          \
      CALL;                          \
    } break;                         \
    case 16: {                       \
      constexpr int DP = 16;         \
      CALL;                          \
    } break;                         \
    default: {                       \
      constexpr int DP = 32;         \
      CALL;                          \
    } break;                         \
  }

}  // namespace

HessSlots hess_slots(const HostProgram& hp) {
  HessSlots H{};
  for (int s = 0; s < hp.ns; ++s) {
    const int l = hp.slot_leaf[s];
    H.po[s] = hp.param_off[l];
    switch (hp.kind[l]) {
      case kLeafArd:
        H.cnt[s] = hp.d;
        break;
      case kLeafPeriodic:
        H.cnt[s] = 2;
        break;
      default:
        H.cnt[s] = 1;
        break;
    }
  }
  return H;
}

void dfill(int dp, const DevProgram* prog, const HessSlots& H, DevPoints x, int param, double* D,
           long ldd, hipStream_t s) {
  const int tiles = lower_tiles(x.ld);
  GPRX_HESS_DISPATCH(dp, hipLaunchKernelGGL((k_dfill<DP>), dim3(tiles), dim3(256), 0, s, prog, H,
                                            x.p, x.ld, x.rows, param, D, ldd));
  GPRX_HIP_CHECK(hipGetLastError());
}

void contract2(int dp, const DevProgram* prog, const HessSlots& H, DevPoints x, const double* Ainv,
               long lda, const double* alpha, int pi, int pj, double* partial, hipStream_t s) {
  const int tiles = lower_tiles(x.ld);
  GPRX_HESS_DISPATCH(dp, hipLaunchKernelGGL((k_contract2<DP>), dim3(tiles), dim3(256), 0, s, prog,
                                            H, x.p, x.ld, x.rows, Ainv, lda, alpha, pi, pj,
                                            partial));
  GPRX_HIP_CHECK(hipGetLastError());
}

void dot_lower(const double* A, const double* B, long ld, int n, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_lower, dim3(n), dim3(256), 0, s, A, B, ld, n, partial);
  GPRX_HIP_CHECK(hipGetLastError());
}

void set_diag(double* D, long ld, int n, double v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_diag, dim3((n + 255) / 256), dim3(256), 0, s, D, ld, n, v);
  GPRX_HIP_CHECK(hipGetLastError());
}

}  // namespace gprx
