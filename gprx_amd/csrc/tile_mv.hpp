// Matrix-vector products over lower 128-tiles, shared by the single-matrix
// (dense.hip) and batched (batch.hip) launches: one workgroup reads one tile
// once and writes its row and column partial sums; a second launch adds the
// partials of each output block in a fixed order (the result does not
// depend on scheduling).
#pragma once

#include "dense.hpp"

namespace gprx {

// One tile, streamed in four 32-column slabs: thread (r, h) holds row r of
// the slab's columns h, h + 2, ..., forms its row sums from registers and
// parks the slab in LDS, where thread (c, part) sums 16 rows of column c.
// vr / vc: the vector over the tile's rows / columns.  Row partials (of L v,
// and of A v with the diagonal tile's c <= r) go to out[0..127], column
// partials (L^T v; strict r > c on a diagonal tile for A v) to out[128..],
// the absolute sums of kTmvSymAbs to out[256..] and out[384..].
template <typename T, int MODE>
__device__ __forceinline__ void tile_mv_tile(const T* __restrict__ Mt, long ld,
                                             const double* __restrict__ vrow,
                                             const double* __restrict__ vcol, bool diag,
                                             double* __restrict__ out) {
  constexpr bool kRow = MODE & 1, kCol = MODE & 2, kSym = MODE == kTmvSym || MODE == kTmvSymAbs;
  constexpr bool kAbs = MODE == kTmvSymAbs;
  __shared__ double slab[32][kTile + 1];
  __shared__ double vr[kTile], vc[kTile];
  __shared__ double cred[2][8][32];
  __shared__ double rred[2][2][kTile];
  const int tid = threadIdx.x;
  if (tid < kTile) {
    vr[tid] = vrow[tid];
    vc[tid] = vcol[tid];
  }
  const int r = tid & 127, h = tid >> 7;
  double racc = 0.0, racc2 = 0.0;
  T mv[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) mv[j] = Mt[r + (long)(h + 2 * j) * ld];
#pragma unroll 1
  for (int sb = 0; sb < 4; ++sb) {
    double m[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) m[j] = (double)mv[j];
    if (sb < 3) {
#pragma unroll
      for (int j = 0; j < 16; ++j) mv[j] = Mt[r + (long)((sb + 1) * 32 + h + 2 * j) * ld];
    }
    __syncthreads();  // vr / vc stored; the previous slab's column sums read
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int cc = h + 2 * j, c = sb * 32 + cc;
      const bool in = !diag || c <= r;
      if (kRow) racc += in ? m[j] * vc[c] : 0.0;
      if (kAbs) racc2 += in ? fabs(m[j]) : 0.0;
      if (kCol) slab[cc][r] = m[j];
    }
    if (kCol) {
      __syncthreads();
      const int cc = tid & 31, pt = tid >> 5, c = sb * 32 + cc;
      double a = 0.0, a2 = 0.0;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int rr = pt * 16 + k;
        const bool in = !diag || (kSym ? rr > c : rr >= c);
        const double x = slab[cc][rr];
        a += in ? x * vr[rr] : 0.0;
        if (kAbs) a2 += in ? fabs(x) : 0.0;
      }
      cred[0][pt][cc] = a;
      if (kAbs) cred[1][pt][cc] = a2;
      __syncthreads();
      if (tid < 32) {
        double sum = 0.0, sum2 = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          sum += cred[0][q][tid];
          if (kAbs) sum2 += cred[1][q][tid];
        }
        out[kTile + sb * 32 + tid] = sum;
        if (kAbs) out[3 * kTile + sb * 32 + tid] = sum2;
      }
    }
  }
  if (kRow) {
    rred[0][h][r] = racc;
    if (kAbs) rred[1][h][r] = racc2;
    __syncthreads();
    if (tid < kTile) {
      out[tid] = rred[0][0][tid] + rred[0][1][tid];
      if (kAbs) out[2 * kTile + tid] = rred[1][0][tid] + rred[1][1][tid];
    }
  }
}

// Output row ii of block b of an nt x nt tile matrix whose tile (q, j)
// partials sit at P + (q (q + 1) / 2 + j) * 4 * 128: the row partials of
// tiles (b, 0..b), then the column partials of tiles (b..nt-1, b), each
// list in eight interleaved running sums (fixed order, eight independent
// loads in flight).  acc2 receives the absolute sums when two is set.
__device__ __forceinline__ void tile_mv_sum(const double* __restrict__ P, int nt, int b, int ii,
                                            int mode, bool two, double& acc, double& acc2) {
  acc = 0.0;
  acc2 = 0.0;
  if (mode & 1) {
    const double* R = P + (long)b * (b + 1) / 2 * 4 * kTile;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0}, s2[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int j = 0;
    for (; j + 8 <= b + 1; j += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        s[u] += R[(long)(j + u) * 4 * kTile + ii];
        if (two) s2[u] += R[(long)(j + u) * 4 * kTile + 2 * kTile + ii];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (j + u > b) continue;
      s[u] += R[(long)(j + u) * 4 * kTile + ii];
      if (two) s2[u] += R[(long)(j + u) * 4 * kTile + 2 * kTile + ii];
    }
    acc += ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    acc2 += ((s2[0] + s2[1]) + (s2[2] + s2[3])) + ((s2[4] + s2[5]) + (s2[6] + s2[7]));
  }
  if (mode & 2) {
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0}, s2[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int q = b;
    for (; q + 8 <= nt; q += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double* p = P + ((long)(q + u) * (q + u + 1) / 2 + b) * 4 * kTile;
        s[u] += p[kTile + ii];
        if (two) s2[u] += p[3 * kTile + ii];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (q + u >= nt) continue;
      const double* p = P + ((long)(q + u) * (q + u + 1) / 2 + b) * 4 * kTile;
      s[u] += p[kTile + ii];
      if (two) s2[u] += p[3 * kTile + ii];
    }
    acc += ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    acc2 += ((s2[0] + s2[1]) + (s2[2] + s2[3])) + ((s2[4] + s2[5]) + (s2[6] + s2[7]));
  }
}

}  // namespace gprx
