// The MFMA tile main loop shared by the dense kernels (see dense.hip for
// the lane layouts).  A 256-thread workgroup computes one TT x TT tile
// (TT = 128, or 64 / 32 for launches too small to fill the chip) of op(A) op(B)
// with four waves, staging BK-deep slices of both operands through
// double-buffered LDS fed by a register prefetch Prefetch<T>::P slices deep,
// and reading each slice's MFMA fragments one slice ahead.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace gprx {

// Workgroups per CU the tile kernels are compiled for.
constexpr int kTileOccupancy = 2;

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

// Tile-kernel configuration for a TT x TT tile: K step BK and the LDS row
// stride LS (elements) of a row-contiguous operand, staged as BK rows of TT.
// A k-contiguous operand is staged as TT rows of BK (see kc_offset), so
// both operands are written to LDS with 16-byte stores.  MFMA step s of a
// slice gives lane group g = lane / 16 the k = (4 r + g) VEC + c, s = r VEC
// + c, so one 16-byte read of a k-contiguous row feeds VEC steps; the four
// groups then read row-contiguous rows VEC apart.  Bank model (it accounts
// for every SQ_LDS_BANK_CONFLICT count measured on these kernels): 32 banks
// of 4 bytes, 128 bytes per pass, so a pass serves 32 lanes of a 4-byte
// access, 16 of an 8-byte and 8 of a 16-byte one.  f32 rows VEC = 4 apart
// then need 4 LS = 16 mod 32 dwords; an f64 pass reads one group's 16
// consecutive elements.  Rows stay 16-byte aligned.  BK keeps two 16-byte
// global loads per thread per operand and slice.
template <typename T, int TT = 128>
struct Cfg;
// BK 8 (f64) / 16 (f32) with two workgroups per CU measured fastest on
// gfx950 among BK in {8, 16, 32}: f32 BK 32 (two workgroups per CU, two
// slices prefetched) ties BK 16 (X^T X at n = 16384 11.45 ms against
// 11.52); f64 BK 16 needs more than 256 registers.
template <>
struct Cfg<double, 128> {
  static constexpr int BK = 8;
  static constexpr int LS = 136;
};
template <>
struct Cfg<float, 128> {
  static constexpr int BK = 16;
  static constexpr int LS = 132;
};
template <>
struct Cfg<double, 64> {
  static constexpr int BK = 16;
  static constexpr int LS = 72;
};
template <>
struct Cfg<float, 64> {
  static constexpr int BK = 32;
  static constexpr int LS = 68;
};
// 32 x 32 tiles for launches of a few 128-tiles (one 16 x 16 block per
// wave): the reduction of such a product spreads over more CUs.
template <>
struct Cfg<double, 32> {
  static constexpr int BK = 16;
  static constexpr int LS = 40;
};
template <>
struct Cfg<float, 32> {
  static constexpr int BK = 32;
  static constexpr int LS = 36;
};

// Element offset of 16-byte vector kv (elements kv VEC ..) of row m in a
// k-contiguous slice of rows of BK elements.  KV = BK / VEC vectors make a
// row of 4 KV dwords, so RP = 8 / KV rows share one sweep of the 32 banks;
// XOR-ing kv with (m / RP) mod KV gives the 8 rows of a fragment-read pass
// (same kv) and the RP rows of a store pass (all kv) distinct bank quads.
template <typename T, int BK>
__device__ __forceinline__ int kc_offset(int m, int kv) {
  constexpr int VEC = 16 / (int)sizeof(T), KV = BK / VEC, RP = 8 / KV;
  static_assert(KV >= 1 && KV <= 8 && 8 % KV == 0, "k-contiguous row of 1..8 vectors");
  return m * BK + ((kv ^ ((m / RP) % KV)) * VEC);
}

// f(std::integral_constant<int, 0>) .. f(integral_constant<N - 1>), unrolled.
template <int I, int N, typename F>
__device__ __forceinline__ void static_for_from(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for_from<I + 1, N>(f);
  }
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_from<0, N>(f);
}

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


template <typename T, int TT = 128>
struct TileSmem {
  // one operand slice: BK x LS (row-contiguous) or TT x BK (k-contiguous)
  static constexpr int kElems = Cfg<T, TT>::BK * Cfg<T, TT>::LS > TT * Cfg<T, TT>::BK
                                    ? Cfg<T, TT>::BK * Cfg<T, TT>::LS
                                    : TT * Cfg<T, TT>::BK;
  __attribute__((aligned(16))) T a[2][kElems];
  __attribute__((aligned(16))) T b[2][kElems];
};

// Global-load stages in flight per thread: a slice leaves HBM P + 2
// iterations before its MFMAs run, which hides the load latency of a launch
// too small to have a second workgroup per CU to switch to.  Measured on
// gfx950 at n = 16384: f32 X^T X 11.84 ms at 3 slices, 11.85 at 4, with 4
// faster in the factor (34.4 against 35.1 ms); f64 factor 57.8 ms at 2
// slices against 58.9 at 3 (3 slices and the second fragment set together
// need more than 256 registers and spill).
template <typename T>
struct Prefetch {
  static constexpr int P = 2;
};
template <>
struct Prefetch<float> {
  static constexpr int P = 4;
};

// Instruction order of one slice: its LDS stores, global loads and the
// next slice's fragment reads each issued between two of its NM MFMAs, so
// the LDS work has finished by the time the last MFMA issues and the
// barrier that closes the slice does not wait on it.
template <int NM, int NG, int NW, int NR>
__device__ __forceinline__ void interleave_slice() {
  constexpr int kMfma = 0x008, kVmemRead = 0x020, kDsRead = 0x100, kDsWrite = 0x200;
  static_for<NW>([&](auto) {
    __builtin_amdgcn_sched_group_barrier(kMfma, 1, 0);
    __builtin_amdgcn_sched_group_barrier(kDsWrite, 1, 0);
  });
  static_for<NG>([&](auto) {
    __builtin_amdgcn_sched_group_barrier(kMfma, 1, 0);
    __builtin_amdgcn_sched_group_barrier(kVmemRead, 1, 0);
  });
  static_for<NR>([&](auto) {
    __builtin_amdgcn_sched_group_barrier(kMfma, 1, 0);
    __builtin_amdgcn_sched_group_barrier(kDsRead, 1, 0);
  });
  if constexpr (NM > NG + NW + NR) __builtin_amdgcn_sched_group_barrier(kMfma, NM - NG - NW - NR, 0);
}

// acc += op(A)[ti tile, kb..kb+nk*BK) op(B)[.., tj tile], with Ablk / Bblk
// the tile origins: A "row-contiguous" operand has element (mn, k) at
// P[mn + k * ld], a "k-contiguous" one at P[k + mn * ld].  Four waves in a
// 2 x 2 grid each own a (TT/2)^2 quarter as FB x FB 16 x 16 MFMA blocks.
// The fragments of slice i + 1 are read from LDS while slice i's MFMAs run
// (two fragment register sets), one barrier per slice; against reading a
// slice's fragments just before its MFMAs this took f64 X^T X at n = 16384
// from 23.0 to 22.35 ms and left f32 unchanged.  With krev the slices run
// from the end of the K-range down (the sum's order is then reversed).
template <typename T, bool TA, bool TB, int TT = 128>
__device__ __forceinline__ void tile_mainloop(const T* __restrict__ Ablk, long lda,
                                              const T* __restrict__ Bblk, long ldb, int kb, int nk,
                                              TileSmem<T, TT>& sm,
                                              typename Acc<T>::V (&acc)[TT / 32][TT / 32],
                                              bool krev = false) {
  constexpr int BK = Cfg<T, TT>::BK;
  constexpr int LS = Cfg<T, TT>::LS;
  constexpr int VEC = 16 / sizeof(T);
  constexpr int P = Prefetch<T>::P;
  constexpr int FB = TT / 32, WT = TT / 2;
  using V = typename Vec<T>::V;
  constexpr int NV = TT * BK / VEC / 256;  // 16-byte vectors per thread per operand
  constexpr int KV = BK / VEC;              // 16-byte vectors per k-contiguous row
  static_assert(BK % (4 * VEC) == 0, "a slice feeds whole 16-byte k-vectors to each lane group");
  static_assert(P >= 2, "the prologue fills two register slots");
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = wave_index();
  const int wm = wave & 1, wn = wave >> 1;
  V ra[P][NV], rb[P][NV];
  auto gload = [&](int k0, V(&xa)[NV], V(&xb)[NV]) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int e = tid + 256 * v;
      if (!TA) {
        const int k = e / (TT / VEC), mn = (e % (TT / VEC)) * VEC;
        xa[v] = *reinterpret_cast<const V*>(Ablk + mn + (long)(k0 + k) * lda);
      } else {
        const int mn = e / KV, k = (e % KV) * VEC;
        xa[v] = *reinterpret_cast<const V*>(Ablk + (k0 + k) + (long)mn * lda);
      }
      if (TB) {
        const int k = e / (TT / VEC), mn = (e % (TT / VEC)) * VEC;
        xb[v] = *reinterpret_cast<const V*>(Bblk + mn + (long)(k0 + k) * ldb);
      } else {
        const int mn = e / KV, k = (e % KV) * VEC;
        xb[v] = *reinterpret_cast<const V*>(Bblk + (k0 + k) + (long)mn * ldb);
      }
    }
  };
  auto lstore = [&](int buf, const V(&xa)[NV], const V(&xb)[NV]) {
    T* a = sm.a[buf];
    T* b = sm.b[buf];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int e = tid + 256 * v;
      if (!TA) {
        const int k = e / (TT / VEC), mn = (e % (TT / VEC)) * VEC;
        *reinterpret_cast<V*>(a + k * LS + mn) = xa[v];
      } else {
        *reinterpret_cast<V*>(a + kc_offset<T, BK>(e / KV, e % KV)) = xa[v];
      }
      if (TB) {
        const int k = e / (TT / VEC), mn = (e % (TT / VEC)) * VEC;
        *reinterpret_cast<V*>(b + k * LS + mn) = xb[v];
      } else {
        *reinterpret_cast<V*>(b + kc_offset<T, BK>(e / KV, e % KV)) = xb[v];
      }
    }
  };
  if (nk <= 0) return;
  // fragments of one slice for this lane: step s = r VEC + c of the slice
  constexpr int NS = BK / 4;
  struct Frag {
    T a[NS][FB], b[NS][FB];
  };
  const int g = lane >> 4, l16 = lane & 15;
  auto fread = [&](int buf, Frag& f) {
    const T* a = sm.a[buf];
    const T* b = sm.b[buf];
#pragma unroll
    for (int r = 0; r < BK / (4 * VEC); ++r) {
      const int kv = 4 * r + g;
#pragma unroll
      for (int i = 0; i < FB; ++i) {
        if constexpr (TA) {
          const V x = *reinterpret_cast<const V*>(a + kc_offset<T, BK>(wm * WT + i * 16 + l16, kv));
#pragma unroll
          for (int c = 0; c < VEC; ++c) f.a[r * VEC + c][i] = x[c];
        } else {
#pragma unroll
          for (int c = 0; c < VEC; ++c)
            f.a[r * VEC + c][i] = a[(kv * VEC + c) * LS + wm * WT + i * 16 + l16];
        }
      }
#pragma unroll
      for (int j = 0; j < FB; ++j) {
        if constexpr (!TB) {
          const V x = *reinterpret_cast<const V*>(b + kc_offset<T, BK>(wn * WT + j * 16 + l16, kv));
#pragma unroll
          for (int c = 0; c < VEC; ++c) f.b[r * VEC + c][j] = x[c];
        } else {
#pragma unroll
          for (int c = 0; c < VEC; ++c)
            f.b[r * VEC + c][j] = b[(kv * VEC + c) * LS + wn * WT + j * 16 + l16];
        }
      }
    }
  };
  // LDS fragment reads of one slice per lane
  constexpr int NR = (BK / (4 * VEC)) * FB * ((TA ? 1 : VEC) + (TB ? VEC : 1));
  auto mma = [&](const Frag& f) {
#pragma unroll
    for (int st = 0; st < NS; ++st)
#pragma unroll
      for (int i = 0; i < FB; ++i)
#pragma unroll
        for (int j = 0; j < FB; ++j) acc[i][j] = mfma16(f.b[st][j], f.a[st][i], acc[i][j]);
  };
  // Slice i: loaded into register slot i % P at iteration i - 2 - P, stored
  // to LDS buffer i % 2 at iteration i - 2, read into fragment set i % 2 at
  // iteration i - 1, multiplied at iteration i.  A buffer is overwritten
  // one barrier after its fragments were read, a slot refilled right after
  // its store.  Loads past the end repeat the last slice and are not used.
  constexpr int U = (P % 2 == 0) ? P : 2 * P;
  // K offset of slice i (clamped to the last slice), last slice first when
  // krev
  auto koff = [&](int i) {
    i = min(i, nk - 1);
    return kb + (krev ? nk - 1 - i : i) * BK;
  };
#pragma unroll
  for (int u = 0; u < P; ++u) gload(koff(u), ra[u], rb[u]);
  lstore(0, ra[0], rb[0]);
  gload(koff(P), ra[0], rb[0]);
  __syncthreads();
  Frag fr[2];
  fread(0, fr[0]);
  lstore(1, ra[1], rb[1]);
  gload(koff(P + 1), ra[1], rb[1]);
  // publishes slice 1, and every wave has read slice 0 before iteration 0
  // overwrites buffer 0
  __syncthreads();
  auto body = [&](int i, auto uc) {
    constexpr int u = decltype(uc)::value;
    fread((u + 1) & 1, fr[(u + 1) & 1]);
    lstore(u & 1, ra[(u + 2) % P], rb[(u + 2) % P]);
    gload(koff(i + 2 + P), ra[(u + 2) % P], rb[(u + 2) % P]);
    mma(fr[u & 1]);
    interleave_slice<NS * FB * FB, 2 * NV, 2 * NV, NR>();
    __syncthreads();
  };
  int it = 0;
  for (; it + U <= nk; it += U) {
    static_for<U>([&](auto uc) { body(it + decltype(uc)::value, uc); });
  }
  static_for<U - 1>([&](auto uc) {
    if (it + decltype(uc)::value < nk) body(it + decltype(uc)::value, uc);
  });
}

// Tile of a launch: (ti, tj) and the K-range its triangular operands allow.
struct TileCoord {
  int ti, tj, kb, ke;
};

__device__ __forceinline__ void tile_krange(int K, int krange, TileCoord& t, int tt = kTile) {
  t.kb = 0;
  t.ke = K;
  if (krange & 1) t.kb = max(t.kb, t.ti * tt);
  if (krange & 2) t.kb = max(t.kb, t.tj * tt);
  if (krange & 4) t.ke = min(t.ke, (t.ti + 1) * tt);
  if (krange & 8) t.ke = min(t.ke, (t.tj + 1) * tt);
}

// Tile of this workgroup from a host-built schedule (see tile_schedule in
// dense.hip); false for a padding slot.
__device__ __forceinline__ bool tile_from_schedule(const unsigned* sched, int K, int krange,
                                                   TileCoord& t, int tt = kTile) {
  // Slot 8k + x holds XCD x's k-th tile.  In a batched launch (blockIdx.y =
  // problem) problem z takes XCD (x + z) mod 8's list, so the XCDs whose
  // lists hold a tile more take turns across the batch.
  const unsigned slot = gridDim.y > 1 ? ((blockIdx.x & ~7u) | ((blockIdx.x + blockIdx.y) & 7u))
                                      : blockIdx.x;
  const unsigned v = sched[slot];
  if (v == 0xffffffffu) return false;
  t.ti = (int)(v >> 16);
  t.tj = (int)(v & 0xffffu);
  tile_krange(K, krange, t, tt);
  return true;
}

__device__ __forceinline__ TileCoord tile_coord(int M, int N, int K, int lower, int order,
                                                int krange, int tt = kTile) {
  const int mt = M / tt, nt = N / tt;
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
  tile_krange(K, krange, t, tt);
  return t;
}

}  // namespace gprx
