// Workgroups per CU the tile kernels are compiled for.
constexpr int kTileOccupancy = 2;

// The MFMA tile main loop shared by the dense kernels (see dense.hip for
// the lane layouts).  A 256-thread workgroup computes one 128 x 128 tile of
// op(A) op(B) with four waves of 64 x 64, staging BK-deep slices of both
// operands through double-buffered LDS with a register prefetch of the next
// slice.
#pragma once

#include "common.hpp"

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

// Row of D held in accumulator element q of lane l (see the header).
template <typename T>
__device__ __forceinline__ int drow(int lane, int q);
template <>
__device__ __forceinline__ int drow<double>(int lane, int q) {
  return (lane >> 4) + 4 * q;
}
template <>
__device__ __forceinline__ int drow<float>(int lane, int q) {
  return 4 * (lane >> 4) + q;
}

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
// BK 8 (f64) / 16 (f32) with two workgroups per CU measured fastest on
// gfx950 among BK in {8, 16, 32} and 1-2 workgroups per CU.
template <>
struct Cfg<double> {
  static constexpr int BK = 8;
  static constexpr int LS = 144;
};
template <>
struct Cfg<float> {
  static constexpr int BK = 16;
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


template <typename T>
struct TileSmem {
  __attribute__((aligned(16))) T a[2][Cfg<T>::BK * Cfg<T>::LS];
  __attribute__((aligned(16))) T b[2][Cfg<T>::BK * Cfg<T>::LS];
};

// acc += op(A)[ti tile, kb..kb+nk*BK) op(B)[.., tj tile], with Ablk / Bblk
// the tile origins: A "row-contiguous" operand has element (mn, k) at
// P[mn + k * ld], a "k-contiguous" one at P[k + mn * ld].
template <typename T, bool TA, bool TB>
__device__ __forceinline__ void tile_mainloop(const T* __restrict__ Ablk, long lda,
                                              const T* __restrict__ Bblk, long ldb, int kb, int nk,
                                              TileSmem<T>& sm, typename Acc<T>::V (&acc)[4][4]) {
  constexpr int BK = Cfg<T>::BK;
  constexpr int LS = Cfg<T>::LS;
  constexpr int VEC = 16 / sizeof(T);
  using V = typename Vec<T>::V;
  constexpr int NV = kTile * BK / VEC / 256;  // 16-byte vectors per thread per operand
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  V ra[NV], rb[NV];
  auto gload = [&](int k0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int e = tid + 256 * v;
      if (!TA) {
        const int k = e / (kTile / VEC), mn = (e % (kTile / VEC)) * VEC;
        ra[v] = *reinterpret_cast<const V*>(Ablk + mn + (long)(k0 + k) * lda);
      } else {
        const int mn = e / (BK / VEC), k = (e % (BK / VEC)) * VEC;
        ra[v] = *reinterpret_cast<const V*>(Ablk + (k0 + k) + (long)mn * lda);
      }
      if (TB) {
        const int k = e / (kTile / VEC), mn = (e % (kTile / VEC)) * VEC;
        rb[v] = *reinterpret_cast<const V*>(Bblk + mn + (long)(k0 + k) * ldb);
      } else {
        const int mn = e / (BK / VEC), k = (e % (BK / VEC)) * VEC;
        rb[v] = *reinterpret_cast<const V*>(Bblk + (k0 + k) + (long)mn * ldb);
      }
    }
  };
  auto lstore = [&](int buf) {
    T* a = sm.a[buf];
    T* b = sm.b[buf];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
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
    const T* a = sm.a[cur];
    const T* b = sm.b[cur];
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
}

// Tile of a launch: (ti, tj) and the K-range its triangular operands allow.
struct TileCoord {
  int ti, tj, kb, ke;
};

__device__ __forceinline__ TileCoord tile_coord(int M, int N, int K, int lower, int order,
                                                int krange) {
  const int mt = M / kTile, nt = N / kTile;
  const int total = lower ? mt * (mt + 1) / 2 : mt * nt;
  int bid = xcd_remap(blockIdx.x, total);
  TileCoord t;
  if (order & 1) bid = total - 1 - bid;
  if (lower) {
    tri_index(bid, t.ti, t.tj);
  } else if (order < 2) {
    t.ti = bid % mt;
    t.tj = bid / mt;
  } else {
    t.tj = bid % nt;
    t.ti = bid / nt;
  }
  t.kb = 0;
  t.ke = K;
  if (krange & 1) t.kb = max(t.kb, t.ti * kTile);
  if (krange & 2) t.kb = max(t.kb, t.tj * kTile);
  if (krange & 4) t.ke = min(t.ke, (t.ti + 1) * kTile);
  if (krange & 8) t.ke = min(t.ke, (t.tj + 1) * kTile);
  return t;
}

}  // namespace gprx
