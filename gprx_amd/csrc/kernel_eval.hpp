// Device-side evaluation of a kernel program at one matrix entry, and the
// per-leaf contraction accumulators shared by the fill, contraction and
// fused GEMM-contraction kernels.
#pragma once

#include "common.hpp"

namespace gprx {

// ---------------------------------------------------------------------------
// Device: one entry of the program.
// ---------------------------------------------------------------------------
struct Leaves {
  double v[kMaxLeaves];
  double f0[kMaxLeaves];
  double f1[kMaxLeaves];
};

template <int DP, bool GRAD>
__device__ __forceinline__ double eval_entry(const DevProgram& P, const double (&dx)[DP], Leaves& L) {
  double s = 0.0, r2 = 0.0;
#pragma unroll
  for (int q = 0; q < DP; ++q) {
    const double t = dx[q] * dx[q];
    s += t;
    r2 += t * P.ard_w[q];
  }
#pragma unroll
  for (int l = 0; l < kMaxLeaves; ++l) {
    L.v[l] = 1.0;
    if (GRAD) {
      L.f0[l] = 0.0;
      L.f1[l] = 0.0;
    }
    if (l < P.nl) {
      const int kd = P.kind[l];
      if (kd == kLeafConst) {
        L.v[l] = P.a[l];
        if (GRAD) L.f0[l] = 1.0;
      } else if (kd == kLeafRbf) {
        L.v[l] = exp(-0.5 * s * P.a[l]);
        if (GRAD) L.f0[l] = s;
      } else if (kd == kLeafArd) {
        L.v[l] = exp(-0.5 * r2);
      } else {
        const double r = sqrt(s);
        const double al = r * P.a[l];
        double sn, cs;
        sincos(al, &sn, &cs);
        L.v[l] = exp(-2.0 * sn * sn * P.b[l]);
        if (GRAD) {
          L.f0[l] = sn * sn;
          L.f1[l] = sn * cs * al;
        }
      }
    }
  }
  double k = 0.0;
#pragma unroll
  for (int t = 0; t < kMaxTerms; ++t) {
    if (t < P.nt) {
      double pr = 1.0;
#pragma unroll
      for (int l = 0; l < kMaxLeaves; ++l)
        if ((P.mask[t] >> l) & 1u) pr *= L.v[l];
      k += pr;
    }
  }
  return k;
}

// Adds w * dK/d(leaf features) of one entry to the accumulators.
template <int DP>
__device__ __forceinline__ void accumulate_entry(const DevProgram& P, const double (&dx)[DP],
                                                 const Leaves& L, double w,
                                                 double (&acc)[kMaxLeaves][kAccPerLeaf],
                                                 double (&accd)[DP]) {
  double gard = 0.0;
#pragma unroll
  for (int l = 0; l < kMaxLeaves; ++l) {
    if (l < P.nl) {
      double cl = 0.0;
#pragma unroll
      for (int t = 0; t < kMaxTerms; ++t) {
        if (t < P.nt && ((P.mask[t] >> l) & 1u)) {
          double pr = 1.0;
#pragma unroll
          for (int l2 = 0; l2 < kMaxLeaves; ++l2)
            if (l2 != l && ((P.mask[t] >> l2) & 1u)) pr *= L.v[l2];
          cl += pr;
        }
      }
      const double g = w * cl * L.v[l];
      acc[l][0] += g * L.f0[l];
      acc[l][1] += g * L.f1[l];
      if (l == P.ard_leaf) gard = g;
    }
  }
  if (P.ard_leaf >= 0) {
#pragma unroll
    for (int q = 0; q < DP; ++q) accd[q] += gard * dx[q] * dx[q];
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block reduction of the accumulators into partial[tile * kMaxAcc + k].
template <int DP>
__device__ __forceinline__ void store_partials(double (&acc)[kMaxLeaves][kAccPerLeaf],
                                               double (&accd)[DP], double* red /* [4][kMaxAcc] */,
                                               double* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int l = 0; l < kMaxLeaves; ++l)
#pragma unroll
    for (int c = 0; c < kAccPerLeaf; ++c) {
      const double v = wave_sum(acc[l][c]);
      if (lane == 0) red[wave * kMaxAcc + l * kAccPerLeaf + c] = v;
    }
#pragma unroll
  for (int q = 0; q < kMaxDim; ++q) {
    double v = 0.0;
    if (q < DP) v = wave_sum(accd[q < DP ? q : 0]);
    if (lane == 0) red[wave * kMaxAcc + kMaxLeaves * kAccPerLeaf + q] = v;
  }
  __syncthreads();
  if (threadIdx.x < kMaxAcc) {
    const int k = threadIdx.x;
    out[k] = ((red[k] + red[kMaxAcc + k]) + red[2 * kMaxAcc + k]) + red[3 * kMaxAcc + k];
  }
}

__device__ __forceinline__ void tri_index_f(int bid, int& ti, int& tj) {
  int t = (int)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
  while ((t + 1) * (t + 2) / 2 <= bid) ++t;
  while (t * (t + 1) / 2 > bid) --t;
  ti = t;
  tj = bid - t * (t + 1) / 2;
}

// Loads the DP coordinates of 128 points starting at row r0 into LDS.
template <int DP>
__device__ __forceinline__ void load_points(const double* __restrict__ p, long ld, int r0,
                                            double (*dst)[kTile]) {
  for (int e = threadIdx.x; e < DP * kTile; e += 256) {
    const int q = e / kTile, i = e % kTile;
    dst[q][i] = p[(long)q * ld + r0 + i];
  }
}

}  // namespace gprx
