// The 128 x 128 diagonal block of the recursive factor: Cholesky and
// inverse X = L^-1 in one workgroup, in f64 arithmetic for either storage
// type (the mixed form's f32 blocks are widened on load and X is rounded to
// f32 on store; the log pivots stay f64).
//
// The block is worked in 16-column steps, right-looking, on four waves:
//   wave 0 carries the chain: the diagonal 16 x 16 block's update, its
//     Cholesky factor and inverse X_D in registers (one pivot after another,
//     broadcasts within 16-lane rows by DPP), then the panel block below it;
//   waves 1-3 form the rest of each panel L_i = A_i X_D^T, the trailing
//     update A_uv -= L_u L_v^T, and row i of X once X_D(i) exists:
//     X_ij = -X_D(i) sum_{k=j}^{i-1} L_ik X_kj.
// Every product is a 16 x 16 MFMA block (v_mfma_f64_16x16x4_f64), and each
// wave's share of a step is fixed at compile time (a longest-first plan), so
// all block indices and reduction lengths are constants.  Two barriers per
// step: after the panel, and after the update and the chain.
//
// LDS: A / L column-major at s[c * DS + r] (r >= c); the strictly lower
// blocks of X transposed at s[r * DS + c] (r > c, off the diagonal blocks);
// the eight X_D(i) whole in xd; a 16 x 16 transpose scratch ws.
//
// Lane layouts of the 16 x 16 x 4 MFMA (see dense.hip): the operands of
// k-step t hold M[l % 16][4 t + l / 16]; the result of a product computed
// as C^T holds C[l % 16][l / 16 + 4 q] in element q.  The two agree, so a
// product's result feeds the next product as an operand without leaving
// registers (used for the diagonal update and for X's second product).
#pragma once

#include "common.hpp"
#include "tile_gemm.hpp"

namespace gprx {
namespace diag {

constexpr int DS = 129;             // column stride of s (odd)
constexpr int XDS = 17;             // column stride of an X_D block in xd
constexpr int XDB = 16 * XDS;       // X_D(i)[a][b] at xd[i * XDB + b * XDS + a]
constexpr int WSS = 17;             // column stride of ws

// Dynamic LDS of the kernel in bytes.
constexpr size_t kLdsBytes = (size_t)(kTile * DS + 8 * XDB + 16 * WSS) * sizeof(double) + 16;

__device__ __forceinline__ d4 mma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Lane L of this lane's 16-lane row (DPP row_newbcast).  The source may have
// been written by the instruction just before, and a DPP read needs two wait
// states after a VALU write.
template <int L>
__device__ __forceinline__ double row_bcast(double v) {
  double r;
  asm volatile("s_nop 1\n\tv_mov_b64_dpp %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf"
               : "=v"(r)
               : "v"(v), "n"(L));
  return r;
}

// acc - (lane L's b) * c with the broadcast folded into the FMA, right after
// b's VALU write (on the pivot chain).
template <int L>
__device__ __forceinline__ double fms_bcast_now(double acc, double b, double c) {
  asm volatile("s_nop 1\n\tv_fmac_f64_dpp %0, -%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
               : "+v"(acc)
               : "v"(b), "v"(c), "n"(L));
  return acc;
}

// The same off the chain, threaded through `braid`: the statement takes
// braid and gives it back untouched, so the compiler places it after
// braid's last definition and before its next use.  That interleaves these
// updates with the dependent steps of the next pivot's square root, which
// the in-order issue would otherwise wait on.  b was written at least one
// pivot earlier, so no wait states are needed.
template <int L>
__device__ __forceinline__ void fms_braid(double& acc, double b, double c, double& braid) {
  asm("v_fmac_f64_dpp %0, -%2, %3 row_newbcast:%4 row_mask:0xf bank_mask:0xf ; %1"
      : "+v"(acc), "+v"(braid)
      : "v"(b), "v"(c), "n"(L));
}

// Step jp's updates off the pivot chain, in the order they are next needed:
// X_D column jp+1, A column jp+2, X_D column jp+2, A column jp+3, ...  Item
// n encodes 2 c + (0: column c of A, 1: column c of X_D).
constexpr int pend_item(int jp, int n) {
  return n % 2 == 0 ? 2 * (jp + 1 + n / 2) + 1 : 2 * (jp + 2 + (n - 1) / 2);
}
constexpr int pend_count(int jp) { return jp < 0 ? 0 : 29 - 2 * jp; }

template <int JP, int N0, int N1>
__device__ __forceinline__ void pending(double (&d)[16], double (&x)[16], double& braid) {
  if constexpr (JP >= 0) {
    static_for_from<N0, N1>([&](auto nc) {
      constexpr int code = pend_item(JP, decltype(nc)::value);
      constexpr int c = code >> 1;
      if constexpr (code & 1)
        fms_braid<c>(x[c], d[JP], x[JP], braid);
      else
        fms_braid<c>(d[c], d[JP], d[JP], braid);
    });
  }
}

// Lower-triangle enumeration by rows: t = u (u + 1) / 2 + v, v <= u.
__device__ __forceinline__ void tri_small(int t, int& u, int& v) {
  u = 0;
  while ((u + 1) * (u + 2) / 2 <= t) ++u;
  v = t - u * (u + 1) / 2;
}

// Row r = lane % 16 of the 16 x 16 block at (c0, c0).
__device__ __forceinline__ void load_row16(const double* s, int c0, double (&d)[16]) {
  const int r = threadIdx.x & 15;
#pragma unroll
  for (int k = 0; k < 16; ++k) d[k] = s[(c0 + k) * DS + c0 + r];
}

// Wave 0: Cholesky factor of the 16 x 16 block whose rows are in d (lane r
// = lane % 16 holds row r; entries right of the diagonal collect values no
// other lane reads) and its inverse X_D, whose column r lane r builds in x
// (L x = e_r by the same right-looking sweep: x_c -= L[c][j] x_j).  The four
// 16-lane rows compute the same values; row 0 stores X_D.
//
// Pivot j: p_j -> y = p_j^-1/2 (rsq and two Newton steps: the hardware rsq
// alone is good to about 5e-8) -> column j -> column j + 1 -> p_(j+1).  Step
// j - 1's other updates fill the gaps between y's dependent steps.  A
// non-positive or non-finite pivot only sets *sbad; the block's values are
// then discarded by the caller.
__device__ __forceinline__ void factor16_inv(double* xd, int* sbad, int c0, double (&d)[16]) {
  const int lane = threadIdx.x & 63, r = lane & 15;
  double x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = k == r ? 1.0 : 0.0;
  bool bad = false;
  double p = row_bcast<0>(d[0]);
  static_for<16>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    constexpr int JP = j - 1, N = pend_count(j - 1);
    bad |= !__builtin_amdgcn_class(p, 0x180);  // +normal or +subnormal
    const double mp = -p;
    double y = __builtin_amdgcn_rsq(p);
    pending<JP, 0, N / 7>(d, x, y);
    double t = mp * y;
    pending<JP, N / 7, 2 * N / 7>(d, x, t);
    double e = fma(t, y, 1.0);
    pending<JP, 2 * N / 7, 3 * N / 7>(d, x, e);
    y = fma(0.5 * y, e, y);
    pending<JP, 3 * N / 7, 4 * N / 7>(d, x, y);
    t = mp * y;
    pending<JP, 4 * N / 7, 5 * N / 7>(d, x, t);
    e = fma(t, y, 1.0);
    pending<JP, 5 * N / 7, 6 * N / 7>(d, x, e);
    y = fma(0.5 * y, e, y);
    pending<JP, 6 * N / 7, N>(d, x, y);
    d[j] *= y;
    x[j] *= y;
    if constexpr (j < 15) {
      d[j + 1] = fms_bcast_now<j + 1>(d[j + 1], d[j], d[j]);
      p = row_bcast<j + 1>(d[j + 1]);
    }
  });
  if (lane < 16) {
    // Column r of X_D.  L's diagonal block is read by no later step; its
    // log pivots come from X_D's diagonal (log L_rr = -log X_rr).
    const int jb = c0 / 16;
#pragma unroll
    for (int k = 0; k < 16; ++k) xd[jb * XDB + r * XDS + k] = x[k];
    if (bad && lane == 0) *sbad = 1;
  }
}

// Panel block i of column block jb: L_i,jb = A_i,jb X_D(jb)^T, returned in
// the C^T accumulator layout and stored over A_i,jb.
__device__ __forceinline__ d4 panel_block(double* s, const double* xd, int jb, int i) {
  const int lane = threadIdx.x & 63, g = lane >> 4, l16 = lane & 15;
  const int c0 = 16 * jb;
  d4 acc = d4{};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const double a = xd[jb * XDB + (4 * t + g) * XDS + l16];  // X_D[l16][4t + g]
    const double b = s[(c0 + 4 * t + g) * DS + 16 * i + l16];  // A[16i + l16][c0 + 4t + g]
    acc = mma(a, b, acc);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) s[(c0 + g + 4 * q) * DS + 16 * i + l16] = acc[q];
  return acc;
}

// A_uv -= L_u,jb L_v,jb^T (u >= v; a diagonal block keeps r >= c).
__device__ __forceinline__ void trail_block(double* s, int jb, int u, int v) {
  const int lane = threadIdx.x & 63, g = lane >> 4, l16 = lane & 15;
  const int c0 = 16 * jb;
  double a[4], b[4], old[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    a[t] = s[(c0 + 4 * t + g) * DS + 16 * v + l16];
    b[t] = s[(c0 + 4 * t + g) * DS + 16 * u + l16];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) old[q] = s[(16 * v + g + 4 * q) * DS + 16 * u + l16];
  d4 acc = d4{};
#pragma unroll
  for (int t = 0; t < 4; ++t) acc = mma(a[t], b[t], acc);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int c = g + 4 * q;
    if (u != v || l16 >= c) s[(16 * v + c) * DS + 16 * u + l16] = old[q] - acc[q];
  }
}

// Block (I, J), J < I, of X: T = sum_{K=J}^{I-1} L_IK X_KJ, then X_IJ =
// -X_D(I) T, both computed as C (not C^T), so T's accumulator is the second
// product's operand.  Stored transposed in s and to the global X.
template <typename T, int I, int J>
__device__ __forceinline__ void xrow_block(double* s, const double* xd, T* __restrict__ X,
                                           long ldx) {
  const int lane = threadIdx.x & 63, g = lane >> 4, l16 = lane & 15;
  d4 acc = d4{};
  static_for_from<J, I>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int m = 16 * K + 4 * t + g;
      const double a = s[m * DS + 16 * I + l16];  // L[16I + l16][m]
      // X[m][16J + l16]: X_D(J) on the diagonal block, the transposed
      // storage below it
      const double b = K == J ? xd[J * XDB + l16 * XDS + 4 * t + g] : s[m * DS + 16 * J + l16];
      acc = mma(a, b, acc);
    }
  });
  d4 res = d4{};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const double a = xd[I * XDB + (4 * t + g) * XDS + l16];  // X_D(I)[l16][4t + g]
    res = mma(a, acc[t], res);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = g + 4 * q;
    const double v = -res[q];
    s[(16 * I + r) * DS + 16 * J + l16] = v;
    X[(16 * I + r) + (long)(16 * J + l16) * ldx] = (T)v;
  }
}

// X_D(i), with the zeros above its diagonal, to the global X, and log L_rr
// of the block's pivots.
template <typename T>
__device__ __forceinline__ void store_xdiag(const double* xd, int i, T* __restrict__ X, long ldx,
                                            double* __restrict__ logdiag) {
  const int lane = threadIdx.x & 63, g = lane >> 4, l16 = lane & 15;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int c = g + 4 * q;
    X[(16 * i + l16) + (long)(16 * i + c) * ldx] = (T)xd[i * XDB + c * XDS + l16];
  }
  if (lane < 16) logdiag[16 * i + l16] = -log(xd[i * XDB + l16 * (XDS + 1)]);
}

// Workgroup barrier that waits for this wave's LDS traffic only: the
// kernel's global stores (X, log pivots) are read by none of its waves.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Step JB's work beside the chain: blocks j of X row JB (cost JB - j + 1
// products of depth 16), the trailing blocks (u, v) of the update (cost 1)
// and the store of X_D(JB) (cost 0).  Waves 1-3 share it (all four at the
// last step, which has no chain); each task goes to the least-loaded wave,
// costliest first.
struct Plan {
  int n;
  int owner[32], u[32], v[32];
};
constexpr Plan make_plan(int jb) {
  Plan p{};
  int load[4] = {0, 0, 0, 0};
  const int w0 = jb == 7 ? 0 : 1;
  int n = 0;
  for (int t = 0; t < 32; ++t) {
    int cost = 0, u = -1, v = -1;
    if (t < jb) {
      cost = jb - t + 1;
      u = jb;
      v = t;
    } else {
      int k = t - jb, found = 0;
      for (int vv = jb + 1; vv < 8 && !found; ++vv)
        for (int uu = vv; uu < 8 && !found; ++uu) {
          if (uu == jb + 1 && vv == jb + 1) continue;
          if (k == 0) {
            u = uu;
            v = vv;
            found = 1;
          } else {
            --k;
          }
        }
      cost = found ? 1 : 0;  // not found: the store of X_D(jb)
    }
    int best = w0;
    for (int w = w0; w < 4; ++w)
      if (load[w] < load[best]) best = w;
    load[best] += cost;
    p.owner[n] = best;
    p.u[n] = u;
    p.v[n] = v;
    ++n;
    if (u < 0) break;
  }
  p.n = n;
  return p;
}
template <int JB>
struct PlanOf {
  static constexpr Plan p = make_plan(JB);
};

template <typename T, int JB, int W>
__device__ __forceinline__ void side_tasks(double* s, const double* xd, T* __restrict__ X,
                                           long ldx, double* __restrict__ logdiag) {
  static_for<PlanOf<JB>::p.n>([&](auto tc) {
    constexpr int t = decltype(tc)::value;
    if constexpr (PlanOf<JB>::p.owner[t] == W) {
      constexpr int u = PlanOf<JB>::p.u[t], v = PlanOf<JB>::p.v[t];
      if constexpr (u < 0)
        store_xdiag<T>(xd, JB, X, ldx, logdiag);
      else if constexpr (t < JB)
        xrow_block<T, JB, v>(s, xd, X, ldx);
      else
        trail_block(s, JB, u, v);
    }
  });
}

// Panel blocks JB + 2 .. 7 on waves 1-3 (block JB + 1 is wave 0's).
template <int JB, int W>
__device__ __forceinline__ void panel_tasks(double* s, const double* xd) {
  static_for_from<JB + 2, 8>([&](auto ic) {
    constexpr int i = decltype(ic)::value;
    if constexpr (1 + (i - JB - 2) % 3 == W) (void)panel_block(s, xd, JB, i);
  });
}

// A: the block (lower part read; the whole 128 x 128 lies inside the
// matrix), X: L^-1 with zeros above the diagonal, logdiag[0..128): log L_rr;
// *info = 1 on a non-positive or non-finite pivot.  A and X must not
// overlap, and both must be 16-byte aligned with lda a multiple of 4.
template <typename T>
__device__ __forceinline__ void factor_block(const T* __restrict__ A, long lda, T* __restrict__ X,
                                             long ldx, double* __restrict__ logdiag, int* info,
                                             double* smem) {
  double* s = smem;
  double* xd = s + kTile * DS;
  double* ws = xd + 8 * XDB;
  int* sbad = reinterpret_cast<int*>(ws + 16 * WSS);
  const int tid = threadIdx.x, lane = tid & 63, wave = wave_index();
  const int g = lane >> 4, l16 = lane & 15;
  if (tid == 0) *sbad = 0;
  if (wave == 0) {
    // block (0, 0) first, so its factor overlaps the other waves' loads
    T v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = A[l16 + (long)(g + 4 * q) * lda];
#pragma unroll
    for (int q = 0; q < 4; ++q) s[(g + 4 * q) * DS + l16] = (double)v[q];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    double d[16];
    load_row16(s, 0, d);
    factor16_inv(xd, sbad, 0, d);
  } else {
    // The rest of the lower triangle in 16-byte vectors (RV rows of one
    // column), all loads in flight.  Vectors wholly above the diagonal are
    // skipped: that part of s holds X, written before it is read.
    using V = typename Vec<T>::V;
    constexpr int RV = 16 / (int)sizeof(T);
    constexpr int NRV = kTile / RV;           // row vectors per column
    constexpr int CS = 192 / NRV;             // column phases
    constexpr int NI = (kTile + CS - 1) / CS;
    const int t3 = tid - 64, r0 = RV * (t3 % NRV), cp = t3 / NRV;
    V v[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int c = cp + CS * i;
      if (c < kTile && r0 + RV - 1 >= c && (c >= 16 || r0 >= 16))
        v[i] = *reinterpret_cast<const V*>(A + r0 + (long)c * lda);
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int c = cp + CS * i;
      if (c < kTile && r0 + RV - 1 >= c && (c >= 16 || r0 >= 16)) {
#pragma unroll
        for (int e = 0; e < RV; ++e) s[c * DS + r0 + e] = (double)v[i][e];
      }
    }
    // zeros of X above the 16-block diagonal (X_D's stores fill the
    // diagonal blocks' upper parts)
    for (int b = wave - 1; b < 28; b += 3) {
      int u, w;
      tri_small(b, u, w);
      const int bi = w, bj = u + 1;
#pragma unroll
      for (int q = 0; q < 4; ++q) X[(16 * bi + l16) + (long)(16 * bj + g + 4 * q) * ldx] = (T)0;
    }
  }
  lds_barrier();
  static_for<8>([&](auto jc) {
    constexpr int JB = decltype(jc)::value;
    d4 pj = d4{};
    double dn[16];
    if constexpr (JB < 7) {
      // panel column JB; wave 0 takes block JB + 1, which the chain needs
      // next, and reads that diagonal block's rows (final since the last
      // step) ahead of its update
      switch (wave) {
        case 0:
          load_row16(s, 16 * (JB + 1), dn);
          pj = panel_block(s, xd, JB, JB + 1);
          break;
        case 1: panel_tasks<JB, 1>(s, xd); break;
        case 2: panel_tasks<JB, 2>(s, xd); break;
        default: panel_tasks<JB, 3>(s, xd); break;
      }
      lds_barrier();
    }
    switch (wave) {
      case 0:
        if constexpr (JB < 7) {
          // A_(JB+1, JB+1) -= P P^T with P the panel block in registers,
          // through ws into the row layout, then the chain
          constexpr int o = 16 * (JB + 1);
          d4 acc = d4{};
#pragma unroll
          for (int t = 0; t < 4; ++t) acc = mma(pj[t], pj[t], acc);
#pragma unroll
          for (int q = 0; q < 4; ++q) ws[(g + 4 * q) * WSS + l16] = acc[q];
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
          for (int k = 0; k < 16; ++k) dn[k] -= ws[k * WSS + l16];
          factor16_inv(xd, sbad, o, dn);
        } else {
          side_tasks<T, JB, 0>(s, xd, X, ldx, logdiag);
        }
        break;
      case 1: side_tasks<T, JB, 1>(s, xd, X, ldx, logdiag); break;
      case 2: side_tasks<T, JB, 2>(s, xd, X, ldx, logdiag); break;
      default: side_tasks<T, JB, 3>(s, xd, X, ldx, logdiag); break;
    }
    lds_barrier();
  });
  if (tid == 0 && *sbad) atomicExch(info, 1);
}

}  // namespace diag
}  // namespace gprx
