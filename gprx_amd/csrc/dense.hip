// Dense linear-algebra kernels for gfx950 (CDNA4).
//
// Everything here works on 128 x 128 tiles of column-major matrices whose
// orders are padded to multiples of 128, so the hot loops carry no bounds
// checks.  Products run on the matrix cores through v_mfma_f64_16x16x4_f64
// (f64) and v_mfma_f32_16x16x4_f32 (f32); both share one lane layout:
//   srcA (16x4):  lane l holds X[l % 16][l / 16]
//   srcB (4x16):  lane l holds Y[l / 16][l % 16]
//   D (16x16):    lane l holds D[r(l, q)][l % 16], q = 0..3, with
//                 r = 4 * (l / 16) + q for f32 and r = l / 16 + 4 * q for f64
//                 (measured on gfx950 with an in-bounds probe of both builtins)
// The tile kernel feeds op(B) as srcA and op(A) as srcB, so D holds C^T and
// a lane's four results sit in four columns of C at one row: consecutive
// lanes store consecutive rows, which is the contiguous direction of a
// column-major C.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "dense.hpp"
#include "diag_block.hpp"
#include "tile_gemm.hpp"
#include "tile_mv.hpp"

namespace gprx {

template <typename T, bool TA, bool TB, int TT>
__global__ __launch_bounds__(256, kTileOccupancy) void k_gemm(GemmArgs<T> g) {
  using AV = typename Acc<T>::V;
  constexpr int FB = TT / 32, WT = TT / 2;
  __shared__ TileSmem<T, TT> sm;
  const int z = blockIdx.y;
  const T* __restrict__ A = g.A + (long)z * g.strideA;
  const T* __restrict__ B = g.B + (long)z * g.strideB;
  T* __restrict__ C = g.C + (long)z * g.strideC;
  TileCoord tc;
  if (g.sched) {
    if (!tile_from_schedule(g.sched, g.K, g.krange, tc, TT)) return;
  } else {
    tc = tile_coord(g.M, g.N, g.K, g.lower, g.order, g.krange, TT);
  }
  const int ti = tc.ti, tj = tc.tj;
  const int nk = tc.ke > tc.kb ? (tc.ke - tc.kb) / Cfg<T, TT>::BK : 0;
  const int lane = threadIdx.x & 63, wave = wave_index();
  const int wm = wave & 1, wn = wave >> 1;
  AV acc[FB][FB];
#pragma unroll
  for (int i = 0; i < FB; ++i)
#pragma unroll
    for (int j = 0; j < FB; ++j) acc[i][j] = AV{};
  const T* Ablk = TA ? A + (long)ti * TT * g.lda : A + (long)ti * TT;
  const T* Bblk = TB ? B + (long)tj * TT : B + (long)tj * TT * g.ldb;
  // K-ranges that start at the tile's row or column but share their end:
  // run them from the end down, so the tiles an XCD starts together read
  // the same operand rows at the same time and share them in its L2 (X^T X
  // at n = 16384: HBM-side reads -18% in f64, -22% in f32; f32 12.18 to
  // 11.77 ms).
  const bool krev = (g.krange & (kKloRow | kKloCol)) && !(g.krange & (kKhiRow | kKhiCol));
  tile_mainloop<T, TA, TB, TT>(Ablk, g.lda, Bblk, g.ldb, tc.kb, nk, sm, acc, krev);

  // Epilogue.  With beta != 0 the old values of a row of FB blocks are read
  // before any of them is stored (the tile lies inside C, so entries above
  // a diagonal tile's diagonal are read and dropped): interleaved, each
  // store could alias the next load and the reads would run one at a time.
  const bool diag_tile = g.lower && ti == tj;
  T* Ct = C + (long)(tj * TT) * g.ldc + ti * TT;
#pragma unroll
  for (int i = 0; i < FB; ++i) {
    T old[FB][4];
    if (g.beta != 0.0) {
#pragma unroll
      for (int j = 0; j < FB; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = wm * WT + i * 16 + (lane & 15);
          const int c = wn * WT + j * 16 + drow<T>(lane, q);
          old[j][q] = Ct[(long)c * g.ldc + r];
        }
    }
#pragma unroll
    for (int j = 0; j < FB; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = wm * WT + i * 16 + (lane & 15);
        const int c = wn * WT + j * 16 + drow<T>(lane, q);
        if (diag_tile && r < c) continue;
        double v = g.alpha * (double)acc[i][j][q];
        if (g.beta != 0.0) v += g.beta * (double)old[j][q];
        Ct[(long)c * g.ldc + r] = (T)v;
      }
  }
}

// ---------------------------------------------------------------------------
// Tile schedules.  The hardware hands workgroup b to XCD b % 8, each XCD
// with its own L2, and within an XCD starts workgroups in order as slots
// free up.  A launch whose operands are triangular gives its tiles K-ranges
// that differ by up to the whole K, so a fixed blockIdx -> tile map that
// keeps neighbours together (for L2 reuse) puts most of the work on a few
// XCDs.  The schedule built here:
//   - groups tiles into S x S super-blocks (S = 2; 1 when the launch has
//     few tiles), so tiles running together on an XCD share 2S operand
//     panels.  Balance counts for more than that sharing: measured on gfx950
//     (A/B runs on one box), S = 2 against 8 took the n = 16384 FP64
//     evaluation from 75.7 to 75.2 ms and the Sgpr FP64 evaluation (n =
//     2^20, m = 1024) from 106.5 to 102.5 ms; S = 16 lost 1.5 ms on the
//     exact evaluation, S = 1 tied S = 2 there and lost 1 ms on Sgpr;
//   - deals super-blocks, heaviest first, each to the least-loaded XCD
//     (longest-processing-time), which balances the XCDs' total work and
//     starts the longest tiles first;
//   - writes XCD x's k-th tile to slot 8k + x, padding with kNoTile.
// Work of a tile is its K-range in 128-blocks plus one for the epilogue.
// Schedules are cached per device and launch shape for the process.
// ---------------------------------------------------------------------------
namespace {

constexpr int kXcds = 8;

struct SchedKey {
  int dev, mt, nt, lower, krange, kt;
  bool operator<(const SchedKey& o) const {
    return std::tie(dev, mt, nt, lower, krange, kt) <
           std::tie(o.dev, o.mt, o.nt, o.lower, o.krange, o.kt);
  }
};

struct Sched {
  unsigned* dev = nullptr;
  int slots = 0;
};

std::mutex g_sched_mu;
std::map<SchedKey, Sched> g_sched;

int tile_work(int ti, int tj, int kt, int krange) {
  int kb = 0, ke = kt;
  if (krange & kKloRow) kb = std::max(kb, ti);
  if (krange & kKloCol) kb = std::max(kb, tj);
  if (krange & kKhiRow) ke = std::min(ke, ti + 1);
  if (krange & kKhiCol) ke = std::min(ke, tj + 1);
  return std::max(0, ke - kb) + 1;
}

}  // namespace

static std::vector<unsigned> build_tile_schedule(int mt, int nt, int lower, int krange, int kt) {
  struct Block {
    long work = 0;
    int order = 0;
    std::vector<unsigned> tiles;
  };
  auto make_blocks = [&](int S) {
    std::vector<Block> blocks;
    const int bm = (mt + S - 1) / S, bn = (nt + S - 1) / S;
    for (int I = 0; I < bm; ++I)
      for (int J = 0; J < bn; ++J) {
        if (lower && J > I) continue;
        Block b;
        b.order = (int)blocks.size();
        for (int tj = J * S; tj < std::min(nt, (J + 1) * S); ++tj)
          for (int ti = I * S; ti < std::min(mt, (I + 1) * S); ++ti) {
            if (lower && tj > ti) continue;
            b.tiles.push_back((unsigned)ti << 16 | (unsigned)tj);
            b.work += tile_work(ti, tj, kt, krange);
          }
        if (!b.tiles.empty()) blocks.push_back(std::move(b));
      }
    return blocks;
  };
  int S = 2;
  std::vector<Block> blocks = make_blocks(S);
  while (S > 1 && (int)blocks.size() < 4 * kXcds) {
    S /= 2;
    blocks = make_blocks(S);
  }
  std::stable_sort(blocks.begin(), blocks.end(),
                   [](const Block& a, const Block& b) { return a.work > b.work; });
  std::vector<std::vector<unsigned>> lists(kXcds);
  long load[kXcds] = {0};
  for (const Block& b : blocks) {
    int x = 0;
    for (int i = 1; i < kXcds; ++i)
      if (load[i] < load[x]) x = i;
    load[x] += b.work;
    lists[x].insert(lists[x].end(), b.tiles.begin(), b.tiles.end());
  }
  size_t len = 0;
  for (auto& l : lists) len = std::max(len, l.size());
  std::vector<unsigned> out(len * kXcds, kNoTile);
  for (int x = 0; x < kXcds; ++x)
    for (size_t k = 0; k < lists[x].size(); ++k) out[k * kXcds + x] = lists[x][k];
  return out;
}

// Device copy of the schedule, or nullptr (the kernel then maps blockIdx.x
// itself) when the stream is being captured and the schedule is not cached:
// the upload is a synchronous copy.
static const unsigned* tile_schedule(int mt, int nt, int lower, int krange, int kt,
                                     hipStream_t stream, int* slots) {
  int dev = 0;
  GPRX_HIP_CHECK(hipGetDevice(&dev));
  const SchedKey key{dev, mt, nt, lower, krange, kt};
  std::lock_guard<std::mutex> lock(g_sched_mu);
  auto it = g_sched.find(key);
  if (it != g_sched.end()) {
    *slots = it->second.slots;
    return it->second.dev;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  GPRX_HIP_CHECK(hipStreamIsCapturing(stream, &cs));
  if (cs != hipStreamCaptureStatusNone) return nullptr;
  const std::vector<unsigned> h = build_tile_schedule(mt, nt, lower, krange, kt);
  Sched sc;
  sc.slots = (int)h.size();
  GPRX_HIP_CHECK(hipMalloc(&sc.dev, sizeof(unsigned) * h.size()));
  GPRX_HIP_CHECK(hipMemcpy(sc.dev, h.data(), sizeof(unsigned) * h.size(), hipMemcpyHostToDevice));
  g_sched[key] = sc;
  *slots = sc.slots;
  return sc.dev;
}

template <typename T, int TT>
static void launch_gemm(bool ta, bool tb, GemmArgs<T> g, hipStream_t stream, bool scheduled) {
  const int mt = g.M / TT, nt = g.N / TT;
  const int total = g.lower ? mt * (mt + 1) / 2 : mt * nt;
  int slots = total;
  g.sched = scheduled ? tile_schedule(mt, nt, g.lower, g.krange, g.K / TT, stream, &slots) : nullptr;
  if (!g.sched) slots = total;
  dim3 grid(slots, g.batch > 0 ? g.batch : 1);
  if (!ta && !tb)
    hipLaunchKernelGGL((k_gemm<T, false, false, TT>), grid, dim3(256), 0, stream, g);
  else if (!ta && tb)
    hipLaunchKernelGGL((k_gemm<T, false, true, TT>), grid, dim3(256), 0, stream, g);
  else if (ta && !tb)
    hipLaunchKernelGGL((k_gemm<T, true, false, TT>), grid, dim3(256), 0, stream, g);
  else
    hipLaunchKernelGGL((k_gemm<T, true, true, TT>), grid, dim3(256), 0, stream, g);
  GPRX_HIP_CHECK(hipGetLastError());
}

// Tile size of a launch, by its number of 128-tiles:
//   - below 256 (fewer than one per CU): 32-tiles, sixteen times the
//     workgroups, so the product spreads over the chip instead of a few CUs;
//   - below 1024 with a triangular operand: 64-tiles.  Its tiles' K-ranges
//     differ by up to the whole K, so with few tiles per CU the longest sets
//     the time; a 64-tile holds a quarter of that work and four times the
//     tiles balance;
//   - otherwise 128-tiles.
// Measured on gfx950 (A/B runs on one box): 32-tiles up to 256 128-tiles
// against 64 took the Svgp step (m = 2048, b = 4096) from 3.06 to 2.98 ms
// mixed and the n = 16384 mixed evaluation from 42.4 to 42.0 ms; the
// triangular rule took the Svgp step from 3.41 to 3.07 ms (mixed) and 4.56
// to 4.03 ms (FP64), whose m = 2048 products X^T H X had run 265 us each on
// 256 128-tiles, and the n = 16384 FP64 factor from 54.8 to 53.5 ms; its
// bound at 512 gave 3.17 / 4.30 ms on the Svgp step, at 2048 it lost 1 ms
// on the exact factor.
constexpr int kTinyTileBelow = 256;
constexpr int kTriSmallTileBelow = 1024;

template <typename T>
static void gemm_tiles(bool ta, bool tb, const GemmArgs<T>& g, hipStream_t stream, bool scheduled) {
  const int mt = g.M / kTile, nt = g.N / kTile;
  const long total = (g.lower ? (long)mt * (mt + 1) / 2 : (long)mt * nt) * (g.batch > 0 ? g.batch : 1);
  if (total == 0) return;
  if (total < kTinyTileBelow)
    launch_gemm<T, 32>(ta, tb, g, stream, scheduled);
  else if (g.krange && total < kTriSmallTileBelow)
    launch_gemm<T, 64>(ta, tb, g, stream, scheduled);
  else
    launch_gemm<T, 128>(ta, tb, g, stream, scheduled);
}

template <typename T>
void gemm(bool ta, bool tb, const GemmArgs<T>& g, hipStream_t stream) {
  gemm_tiles<T>(ta, tb, g, stream, true);
}

template <typename T>
void gemm_unscheduled(bool ta, bool tb, const GemmArgs<T>& g, hipStream_t stream) {
  gemm_tiles<T>(ta, tb, g, stream, false);
}

template void gemm<double>(bool, bool, const GemmArgs<double>&, hipStream_t);
template void gemm<float>(bool, bool, const GemmArgs<float>&, hipStream_t);
template void gemm_unscheduled<double>(bool, bool, const GemmArgs<double>&, hipStream_t);
template void gemm_unscheduled<float>(bool, bool, const GemmArgs<float>&, hipStream_t);

// ---------------------------------------------------------------------------
// 128 x 128 diagonal blocks (diag_block.hpp): one workgroup per block.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_potrf_diag(const T* A, long lda, T* X, long ldx,
                                                    double* logdiag, int* info, long sa, long sx,
                                                    long slog) {
  extern __shared__ double smem[];
  const long z = blockIdx.x;
  diag::factor_block<T>(A + z * sa, lda, X + z * sx, ldx, logdiag + z * slog, info, smem);
}

template <typename T>
__global__ __launch_bounds__(256) void k_potrf_diag_list(T* A, T* X, const long* a_off,
                                                         const long* x_off, const long* ld,
                                                         const long* log_off, double* logdiag,
                                                         int* info, const int* info_idx) {
  extern __shared__ double smem[];
  const int b = blockIdx.x;
  diag::factor_block<T>(A + a_off[b], ld[b], X + x_off[b], ld[b], logdiag + log_off[b],
                        info + info_idx[b], smem);
}

template <typename T>
constexpr size_t kDiagLds = diag::kLdsBytes;

// The dynamic-LDS limit is raised once per device (the kernels need more
// than the default 64 KiB); callers on several host threads share the flags.
template <typename T>
static void set_diag_lds_once() {
  static std::mutex mu;
  static std::vector<char> done;
  int dev = 0;
  GPRX_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if ((int)done.size() <= dev) done.resize(dev + 1, 0);
  if (done[dev]) return;
  GPRX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_potrf_diag<T>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDiagLds<T>));
  GPRX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_potrf_diag_list<T>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDiagLds<T>));
  done[dev] = 1;
}

void prepare_dense_kernels() {
  set_diag_lds_once<double>();
  set_diag_lds_once<float>();
}

template <typename T>
void potrf_diag(const T* A, long lda, T* X, long ldx, double* logdiag, int* info, int batch,
                long stride_a, long stride_x, long stride_log, hipStream_t stream) {
  set_diag_lds_once<T>();
  hipLaunchKernelGGL((k_potrf_diag<T>), dim3(batch), dim3(256), kDiagLds<T>, stream, A, lda, X, ldx,
                     logdiag, info, stride_a, stride_x, stride_log);
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T>
void potrf_diag_list(T* A, T* X, const long* a_off, const long* x_off, const long* ld,
                     const long* log_off, double* logdiag, int* info, const int* info_idx, int count,
                     hipStream_t stream) {
  if (count <= 0) return;
  set_diag_lds_once<T>();
  hipLaunchKernelGGL((k_potrf_diag_list<T>), dim3(count), dim3(256), kDiagLds<T>, stream, A, X, a_off,
                     x_off, ld, log_off, logdiag, info, info_idx);
  GPRX_HIP_CHECK(hipGetLastError());
}

template void potrf_diag<double>(const double*, long, double*, long, double*, int*, int, long, long,
                                 long, hipStream_t);
template void potrf_diag<float>(const float*, long, float*, long, double*, int*, int, long, long,
                                long, hipStream_t);
template void potrf_diag_list<double>(double*, double*, const long*, const long*, const long*,
                                      const long*, double*, int*, const int*, int, hipStream_t);
template void potrf_diag_list<float>(float*, float*, const long*, const long*, const long*,
                                     const long*, double*, int*, const int*, int, hipStream_t);

// ---------------------------------------------------------------------------
// Matrix-vector products (HBM-bound; f64 accumulation, fixed order).
// ---------------------------------------------------------------------------
// Columns per split-K chunk.  Wider chunks (so fewer partials per row)
// measured slower: at 1024 x 2^20 (f32) 3.50 ms per product with 4096
// columns per thread against 1.23 ms with 256.
constexpr int kGemvChunk = 256;

template <typename T>
__global__ __launch_bounds__(256) void k_gemv_n_part(const T* __restrict__ M, long ld, int rows,
                                                     int cols, int chunk, const double* __restrict__ v,
                                                     double* __restrict__ part, int tri, int strict,
                                                     int abs_mode, int vec) {
  const int c0 = blockIdx.y * chunk;
  if (vec) {
    // RV consecutive rows per thread in one 16-byte load per column (f32:
    // four rows; 4-byte loads had left the f32 product at 3.5 TB/s); each
    // row's sum runs over the same columns in the same order
    constexpr int RV = 16 / (int)sizeof(T);
    using V = typename Vec<T>::V;
    const int i0 = (blockIdx.x * 256 + threadIdx.x) * RV;
    if (i0 >= rows) return;
    const int c1 = min(cols, c0 + chunk);
    double acc[RV];
#pragma unroll
    for (int e = 0; e < RV; ++e) acc[e] = 0.0;
    for (int c = c0; c < c1; ++c) {
      const V mv = *reinterpret_cast<const V*>(M + i0 + (long)c * ld);
      const double vc = v[c];
#pragma unroll
      for (int e = 0; e < RV; ++e) {
        const double m = (double)mv[e];
        acc[e] += (abs_mode ? fabs(m) : m) * vc;
      }
    }
#pragma unroll
    for (int e = 0; e < RV; ++e) part[(long)blockIdx.y * rows + i0 + e] = acc[e];
    return;
  }
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= rows) return;
  int c1 = min(cols, c0 + chunk);
  if (tri) c1 = min(c1, strict ? i : i + 1);
  double acc = 0.0;
  for (int c = c0; c < c1; ++c) {
    const double m = (double)M[i + (long)c * ld];
    acc += (abs_mode ? fabs(m) : m) * v[c];
  }
  part[(long)blockIdx.y * rows + i] = acc;
}

// out[i] = sum over the nparts chunk partials of row i: 64 rows per block,
// four threads per row each summing every fourth partial in eight running
// sums, then the four in order (fixed order).
__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ part, int rows, int nparts,
                                                   double* __restrict__ out) {
  __shared__ double red[4][64];
  const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + r;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (i < rows) {
    int p = g;
    for (; p + 28 < nparts; p += 32) {
#pragma unroll
      for (int u = 0; u < 8; ++u) s[u] += part[(long)(p + 4 * u) * rows + i];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (p + 4 * u < nparts) s[u] += part[(long)(p + 4 * u) * rows + i];
  }
  red[g][r] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
  __syncthreads();
  if (g == 0 && i < rows) out[i] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
}

template <typename T>
void gemv_n(const T* M, long ld, int rows, int cols, const double* v, double* out, double* part,
            int tri, int strict, int abs_mode, hipStream_t stream) {
  const int chunk = kGemvChunk;
  const int nparts = (cols + chunk - 1) / chunk;
  constexpr int RV = 16 / (int)sizeof(T);
  // the wide rows only where the launch keeps at least 2048 workgroups (the
  // Svgp step's 2048 x 4096 products have 16 chunks and ran slower wide)
  const long wide_blocks = (long)((rows + 256 * RV - 1) / (256 * RV)) * nparts;
  const int vec = !tri && rows % RV == 0 && ld % RV == 0 &&
                  reinterpret_cast<uintptr_t>(M) % 16 == 0 && wide_blocks >= 2048;
  const int per_block = vec ? 256 * RV : 256;
  dim3 grid((rows + per_block - 1) / per_block, nparts);
  hipLaunchKernelGGL((k_gemv_n_part<T>), grid, dim3(256), 0, stream, M, ld, rows, cols, chunk, v, part,
                     tri, strict, abs_mode, vec);
  hipLaunchKernelGGL(k_sum_parts, dim3((rows + 63) / 64), dim3(256), 0, stream, part, rows, nparts, out);
  GPRX_HIP_CHECK(hipGetLastError());
}

// Columns per wave of k_gemv_t (at most): in the 16-byte path they share
// each lane's loads of v (8 bytes of v per 4 bytes of an f32 column had held
// the f32 product near 3.3 TB/s).  Launches that would then have fewer than
// 2048 workgroups keep one column per wave.  Each column's sum keeps its own
// order.
constexpr int kGemvTCols = 4;

static int gemv_t_cols_per_wave(int cols) { return cols >= 4 * kGemvTCols * 2048 ? kGemvTCols : 1; }

template <typename T>
__global__ __launch_bounds__(256) void k_gemv_t(const T* __restrict__ M, long ld, int rows,
                                                int cols, const double* __restrict__ v,
                                                double* __restrict__ out, int tri, int strict,
                                                int abs_mode, int accumulate, int sumsq, int cpw) {
  const int lane = threadIdx.x & 63;
  const int j0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * cpw;
  if (j0 >= cols) return;
  const int nc = min(cpw, cols - j0);
  constexpr int VEC = 16 / sizeof(T);
  using V = typename Vec<T>::V;
  double acc[kGemvTCols];
#pragma unroll
  for (int c = 0; c < kGemvTCols; ++c) acc[c] = 0.0;
  if (!tri && rows % (64 * VEC) == 0 && ld % VEC == 0 && reinterpret_cast<uintptr_t>(M) % 16 == 0) {
    // whole 16-byte-aligned columns: one 16-byte load per lane, column and step
    for (int i = lane * VEC; i < rows; i += 64 * VEC) {
      double vi[VEC];
#pragma unroll
      for (int e = 0; e < VEC; ++e) vi[e] = sumsq ? 0.0 : v[i + e];
#pragma unroll
      for (int c = 0; c < kGemvTCols; ++c) {
        if (c < nc) {
          const V mv = *reinterpret_cast<const V*>(M + (long)(j0 + c) * ld + i);
#pragma unroll
          for (int e = 0; e < VEC; ++e) {
            const double m = (double)mv[e];
            if (sumsq)
              acc[c] += m * m;
            else
              acc[c] += (abs_mode ? fabs(m) : m) * vi[e];
          }
        }
      }
    }
  } else {
    for (int c = 0; c < nc; ++c) {
      const int j = j0 + c;
      const int r0 = tri ? (strict ? j + 1 : j) : 0;
      const T* col = M + (long)j * ld;
      for (int i = r0 + lane; i < rows; i += 64) {
        const double m = (double)col[i];
        if (sumsq)
          acc[c] += m * m;
        else
          acc[c] += (abs_mode ? fabs(m) : m) * v[i];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < kGemvTCols; ++c) {
    const double a = wave_sum(acc[c]);
    if (lane == 0 && c < nc) out[j0 + c] = accumulate ? out[j0 + c] + a : a;
  }
}

template <typename T>
void gemv_t(const T* M, long ld, int rows, int cols, const double* v, double* out, int tri,
            int strict, int abs_mode, int accumulate, hipStream_t stream) {
  const int cpw = gemv_t_cols_per_wave(cols), per_block = 4 * cpw;
  hipLaunchKernelGGL((k_gemv_t<T>), dim3((cols + per_block - 1) / per_block), dim3(256), 0, stream, M,
                     ld, rows, cols, v, out, tri, strict, abs_mode, accumulate, 0, cpw);
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T>
void col_sumsq(const T* M, long ld, int rows, int cols, double* out, hipStream_t stream) {
  const int cpw = gemv_t_cols_per_wave(cols), per_block = 4 * cpw;
  hipLaunchKernelGGL((k_gemv_t<T>), dim3((cols + per_block - 1) / per_block), dim3(256), 0, stream, M,
                     ld, rows, cols, (const double*)nullptr, out, 0, 0, 0, 0, 1, cpw);
  GPRX_HIP_CHECK(hipGetLastError());
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_tile_mv_lower(const T* __restrict__ M, long ld,
                                                       const double* __restrict__ v,
                                                       double* __restrict__ part) {
  int ti, tj;
  tri_index(blockIdx.x, ti, tj);
  tile_mv_tile<T, MODE>(M + (long)ti * kTile + (long)tj * kTile * ld, ld, v + ti * kTile, v + tj * kTile,
                        ti == tj, part + (long)blockIdx.x * 4 * kTile);
}

__global__ void k_tile_mv_reduce(const double* __restrict__ part, int nt, int mode,
                                 double* __restrict__ out, double* __restrict__ out2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nt * kTile) return;
  double acc, acc2;
  tile_mv_sum(part, nt, i / kTile, i % kTile, mode, out2 != nullptr, acc, acc2);
  out[i] = acc;
  if (out2) out2[i] = acc2;
}

template <typename T>
void tile_mv_lower(const T* M, long ld, int nt, int mode, const double* v, double* out, double* out2,
                   double* part, hipStream_t stream) {
  const int tiles = nt * (nt + 1) / 2;
  switch (mode) {
    case kTmvRow:
      hipLaunchKernelGGL((k_tile_mv_lower<T, kTmvRow>), dim3(tiles), dim3(256), 0, stream, M, ld, v, part);
      break;
    case kTmvCol:
      hipLaunchKernelGGL((k_tile_mv_lower<T, kTmvCol>), dim3(tiles), dim3(256), 0, stream, M, ld, v, part);
      break;
    case kTmvSym:
      hipLaunchKernelGGL((k_tile_mv_lower<T, kTmvSym>), dim3(tiles), dim3(256), 0, stream, M, ld, v, part);
      break;
    case kTmvSymAbs:
      hipLaunchKernelGGL((k_tile_mv_lower<T, kTmvSymAbs>), dim3(tiles), dim3(256), 0, stream, M, ld, v,
                         part);
      break;
    default:
      throw DeviceError("unknown tile product mode");
  }
  hipLaunchKernelGGL(k_tile_mv_reduce, dim3((nt * kTile + 255) / 256), dim3(256), 0, stream, part, nt,
                     mode & 3, out, mode == kTmvSymAbs ? out2 : nullptr);
  GPRX_HIP_CHECK(hipGetLastError());
}

template void tile_mv_lower<double>(const double*, long, int, int, const double*, double*, double*,
                                    double*, hipStream_t);
template void tile_mv_lower<float>(const float*, long, int, int, const double*, double*, double*,
                                   double*, hipStream_t);

template void gemv_n<double>(const double*, long, int, int, const double*, double*, double*, int,
                             int, int, hipStream_t);
template void gemv_n<float>(const float*, long, int, int, const double*, double*, double*, int, int,
                            int, hipStream_t);
template void gemv_t<double>(const double*, long, int, int, const double*, double*, int, int, int,
                             int, hipStream_t);
template void gemv_t<float>(const float*, long, int, int, const double*, double*, int, int, int,
                            int, hipStream_t);
template void col_sumsq<double>(const double*, long, int, int, double*, hipStream_t);
template void col_sumsq<float>(const float*, long, int, int, double*, hipStream_t);

template <typename T>
__global__ void k_zero(T* p, size_t count) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t k = i; k < count; k += stride) p[k] = T(0);
}

template <typename T>
void zero(T* p, size_t count, hipStream_t stream) {
  if (count == 0) return;
  size_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((k_zero<T>), dim3((unsigned)blocks), dim3(256), 0, stream, p, count);
  GPRX_HIP_CHECK(hipGetLastError());
}
template void zero<double>(double*, size_t, hipStream_t);
template void zero<float>(float*, size_t, hipStream_t);

}  // namespace gprx
