// Device-side evaluation of a kernel program at one matrix entry, and the
// per-slot contraction accumulators shared by the fill, contraction and
// fused GEMM-contraction kernels.
#pragma once

#include "common.hpp"

namespace gprx {

// ---------------------------------------------------------------------------
// Device: one entry of the program.
// ---------------------------------------------------------------------------
// Per-entry arithmetic in R (double, or float for the mixed form's f32
// matrices): leaf features and term products; accumulation stays f64.
// NS: slots held (at least the program's ns); fewer slots, fewer registers.
template <typename R, int NS = kMaxSlots>
struct LeavesT {
  R f0[NS];  // d v_s / d log theta = v_s * (feature) per slot
  R f1[NS];
  R tp[NS];  // product of the slot's term up to and including it
};
using Leaves = LeavesT<double>;

__device__ __forceinline__ double k_exp(double x) { return exp(x); }
__device__ __forceinline__ float k_exp(float x) { return expf(x); }
__device__ __forceinline__ double k_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float k_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ void k_sincos(double x, double* s, double* c) { sincos(x, s, c); }
__device__ __forceinline__ void k_sincos(float x, float* s, float* c) { sincosf(x, s, c); }

template <typename R>
__device__ __forceinline__ R prog_a(const DevProgram& P, int s);
template <>
__device__ __forceinline__ double prog_a<double>(const DevProgram& P, int s) { return P.a[s]; }
template <>
__device__ __forceinline__ float prog_a<float>(const DevProgram& P, int s) { return P.af[s]; }
template <typename R>
__device__ __forceinline__ R prog_b(const DevProgram& P, int s);
template <>
__device__ __forceinline__ double prog_b<double>(const DevProgram& P, int s) { return P.b[s]; }
template <>
__device__ __forceinline__ float prog_b<float>(const DevProgram& P, int s) { return P.bf[s]; }
template <typename R>
__device__ __forceinline__ R prog_w(const DevProgram& P, int q);
template <>
__device__ __forceinline__ double prog_w<double>(const DevProgram& P, int q) { return P.ard_w[q]; }
template <>
__device__ __forceinline__ float prog_w<float>(const DevProgram& P, int q) { return P.ard_wf[q]; }

// K at one entry (dx = a - b over the padded dimensions): leaf values slot by
// slot, the running product of the open term, and the sum over closed terms.
template <int DP, bool GRAD, typename R, int NS>
__device__ __forceinline__ R eval_entry(const DevProgram& P, const R (&dx)[DP], LeavesT<R, NS>& L) {
  R s = 0, r2 = 0;
#pragma unroll
  for (int q = 0; q < DP; ++q) {
    const R t = dx[q] * dx[q];
    s += t;
    r2 += t * prog_w<R>(P, q);
  }
  R k = 0, rp = 1;
#pragma unroll
  for (int sl = 0; sl < NS; ++sl) {
    if (sl >= P.ns) break;
    const int kd = P.kind[sl];
    R v, f0 = 0, f1 = 0;
    if (kd == kLeafConst) {
      v = prog_a<R>(P, sl);
      f0 = 1;
    } else if (kd == kLeafRbf) {
      v = k_exp(R(-0.5) * s * prog_a<R>(P, sl));
      f0 = s;
    } else if (kd == kLeafArd) {
      v = k_exp(R(-0.5) * r2);
    } else {
      const R r = k_sqrt(s);
      const R al = r * prog_a<R>(P, sl);
      R sn, cs;
      k_sincos(al, &sn, &cs);
      v = k_exp(R(-2) * sn * sn * prog_b<R>(P, sl));
      f0 = sn * sn;
      f1 = sn * cs * al;
    }
    rp = ((P.start >> sl) & 1u) ? v : rp * v;
    L.tp[sl] = rp;
    if (GRAD) {
      L.f0[sl] = f0;
      L.f1[sl] = f1;
    }
    if ((P.end >> sl) & 1u) k += rp;
  }
  return k;
}

// Adds w * dK/d(slot features) of one entry to the accumulators (type A:
// f64, or f32 for a thread's own short run of f32 entries): sweeping the
// slots backwards, each term's full product (held by its last slot) is
// carried to the slots before it.
template <int DP, typename R, int NS, typename A>
__device__ __forceinline__ void accumulate_entry(const DevProgram& P, const R (&dx)[DP],
                                                 const LeavesT<R, NS>& L, A w,
                                                 A (&acc)[NS][kAccPerLeaf], A (&accd)[DP]) {
  A gard = 0, term = 0;
#pragma unroll
  for (int sl = NS - 1; sl >= 0; --sl) {
    if (sl < P.ns) {
      if ((P.end >> sl) & 1u) term = (A)L.tp[sl];
      const A g = w * term;
      acc[sl][0] += g * (A)L.f0[sl];
      acc[sl][1] += g * (A)L.f1[sl];
      if ((P.ard >> sl) & 1u) gard += g;
    }
  }
  if (P.ard) {
#pragma unroll
    for (int q = 0; q < DP; ++q) accd[q] += gard * (A)(dx[q] * dx[q]);
  }
}

// Workgroups per CU the contraction kernels are compiled for: three (at
// most 168 registers) up to 8 padded dimensions, where it took the Sgpr
// FP64 evaluation (n = 2^20, m = 1024, three leaf slots) from 101.8 to 96.9
// ms; at 16 dimensions (the Svgp step) the bound cost 4%, and four
// workgroups spilled.
// Eight leaf slots get two (256 registers): under three their f64 kernels
// spilled to scratch (up to 160 bytes per lane at 8 padded dimensions).
constexpr int contract_occupancy(int dp, int ns) { return dp <= 8 ? (ns <= 4 ? 3 : 2) : 1; }

// Block reduction of the accumulators (widened to f64) into
// partial[tile * kMaxAcc + k] (slots from NS on hold zeros).
template <int DP, int NS, typename A>
__device__ __forceinline__ void store_partials(A (&acc)[NS][kAccPerLeaf], A (&accd)[DP],
                                               double* red /* [4][kMaxAcc] */, double* out) {
  const int lane = threadIdx.x & 63, wave = wave_index();
#pragma unroll
  for (int l = 0; l < kMaxSlots; ++l)
#pragma unroll
    for (int c = 0; c < kAccPerLeaf; ++c) {
      const double v = l < NS ? wave_sum((double)acc[l < NS ? l : 0][c]) : 0.0;
      if (lane == 0) red[wave * kMaxAcc + l * kAccPerLeaf + c] = v;
    }
#pragma unroll
  for (int q = 0; q < kMaxDim; ++q) {
    double v = 0.0;
    if (q < DP) v = wave_sum((double)accd[q < DP ? q : 0]);
    if (lane == 0) red[wave * kMaxAcc + kMaxSlots * kAccPerLeaf + q] = v;
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
