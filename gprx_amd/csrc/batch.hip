// Many independent exact-GPR problems of uneven n in grouped launches.
//
// Every problem is padded to whole 128-tiles and lives in two arenas (the
// kernel matrix, then L, Y and A^-1; and X = L^-1).  The factor is
// left-looking and level-synchronous: step j of every problem with more than
// j tiles runs in the same three launches,
//   1. column update A_ij -= L_i,0:j L_j,0:j^T, i >= j (one tile task each,
//      reduction length 128 j)
//   2. diagonal block j: Cholesky + inverse in LDS (in f64 for both forms) -> X_jj
//   3. panel L_ij = A_ij X_jj^T (i > j) and Y_jk = X_jj L_jk (k < j)
// then the inverse by distance, X_ij = -sum_k Y_ik X_kj (reduction 128 (i-j)),
// A^-1 = X^T X, alpha and the fused gradient contraction, each as one
// grouped launch.  Task lists are built once; each launch runs a slot
// schedule over the 8 XCDs picked by timing candidates on the batch itself.
// Each evaluation is captured in hipGraphs and replayed per optimizer
// iteration with only the per-problem parameters re-uploaded.  The mixed
// form is two graphs: the first factors, forms alpha and refines it twice
// with f64 residuals on every problem, recording each residual pass; one
// readback applies the refinement's stop rules to them, the few problems
// still refining go on uncaptured over their own tiles, and the second
// graph forms A^-1 and the gradient.  A problem whose f32 factor failed or
// whose refinement did not converge is evaluated alone in FP64.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>

#include "kernel_eval.hpp"
#include "model.hpp"
#include "tile_gemm.hpp"
#include "tile_mv.hpp"

namespace gprx {
int exact_validate_theta(const HostProgram& hp, const double* th, size_t p, gprx_error* err);
}

using namespace gprx;

namespace {

constexpr double kLog2Pi = 1.8378770664093453;

// One output tile of a grouped product: C(c) = beta C + alpha op(A) op(B),
// op(A) 128 x K at a, op(B) K x 128 at b (element offsets into the arenas).
// Each operand lives in arena 0 (the kernel matrix, then L, Y and A^-1) or
// arena 1 (X = L^-1): bit 0 of `arenas` for A, bit 1 for B, bit 2 for C.
struct TileTask {
  long a, b, c;
  int ld;
  int K;
  int tb;     // B stored N x K (NT) instead of K x N (NN)
  int ta;     // A stored K x M (TN)
  int lower;  // bit 0: diagonal tile, keep r >= c; bit 1: run the K slices from the end
  int arenas;
  float alpha, beta;
};

// Sub-tile `quad` (TT = 64: row half quad & 1, column half quad >> 1) of a
// task, or the whole tile (TT = 128).
template <typename T, bool TA, bool TB, int TT>
__device__ __forceinline__ void run_task(const TileTask& t, int quad, T* X0, T* X1,
                                         TileSmem<T, TT>& sm) {
  using AV = typename Acc<T>::V;
  constexpr int FB = TT / 32, WT = TT / 2;
  const int qr = TT == 128 ? 0 : (quad & 1) * 64, qc = TT == 128 ? 0 : (quad >> 1) * 64;
  const T* Ab = ((t.arenas & 1) ? X1 : X0) + t.a + (TA ? (long)qr * t.ld : qr);
  const T* Bb = ((t.arenas & 2) ? X1 : X0) + t.b + (TB ? qc : (long)qc * t.ld);
  T* C = ((t.arenas & 4) ? X1 : X0) + t.c + qr + (long)qc * t.ld;
  AV acc[FB][FB];
#pragma unroll
  for (int i = 0; i < FB; ++i)
#pragma unroll
    for (int j = 0; j < FB; ++j) acc[i][j] = AV{};
  tile_mainloop<T, TA, TB, TT>(Ab, t.ld, Bb, t.ld, 0, t.K / Cfg<T, TT>::BK, sm, acc, (t.lower & 2) != 0);
  const int lane = threadIdx.x & 63, wave = wave_index();
  const int wm = wave & 1, wn = wave >> 1;
  const bool diag = (t.lower & 1) && qr == qc;
  // old values of a row of FB blocks before their stores (as k_gemm's epilogue)
#pragma unroll
  for (int i = 0; i < FB; ++i) {
    T old[FB][4];
    if (t.beta != 0.f) {
#pragma unroll
      for (int j = 0; j < FB; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = wm * WT + i * 16 + (lane & 15);
          const int c = wn * WT + j * 16 + drow<T>(lane, q);
          old[j][q] = C[(long)c * t.ld + r];
        }
    }
#pragma unroll
    for (int j = 0; j < FB; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = wm * WT + i * 16 + (lane & 15);
        const int c = wn * WT + j * 16 + drow<T>(lane, q);
        if (diag && r < c) continue;
        double v = (double)t.alpha * (double)acc[i][j][q];
        if (t.beta != 0.f) v += (double)t.beta * (double)old[j][q];
        C[(long)c * t.ld + r] = (T)v;
      }
  }
}

// One launch of a task range: slot s (blockIdx.x) runs sched[s] = task << 2
// | quadrant, or nothing for a padding slot (-1).  The host builds `sched`
// per launch so XCD x (slots 8k + x) gets a balanced share (see
// build_slots).
template <typename T, int TT>
__global__ __launch_bounds__(256, kTileOccupancy) void k_gemm_tasks(const TileTask* __restrict__ tasks,
                                                                    const int* __restrict__ sched,
                                                                    T* X0, T* X1) {
  __shared__ TileSmem<T, TT> sm;
  const int v = sched[blockIdx.x];
  if (v < 0) return;
  const TileTask t = tasks[v >> 2];
  if (t.ta)
    run_task<T, true, false, TT>(t, v & 3, X0, X1, sm);
  else if (t.tb)
    run_task<T, false, true, TT>(t, v & 3, X0, X1, sm);
  else
    run_task<T, false, false, TT>(t, v & 3, X0, X1, sm);
}

// Per-problem descriptors on the device.
struct ProbDesc {
  long xoff;   // points (rows npad, dim dp) in the point arena
  long voff;   // vectors (npad) in the vector arena
  long moff;   // matrices (npad x npad) in the arenas
  int n, npad;
};

// Lower-tile fill of every problem's K + noise I.
struct FillTask {
  int prob, ti, tj;
};

template <int DP, typename T, bool DUAL>
__global__ __launch_bounds__(256) void k_fill_tasks(const FillTask* __restrict__ tasks,
                                                    const ProbDesc* __restrict__ probs,
                                                    const DevProgram* __restrict__ progs,
                                                    const double* __restrict__ xar, T* Aar,
                                                    double* K64, int* info) {
  __shared__ double xr[DP][kTile];
  __shared__ double xc[DP][kTile];
  const FillTask t = tasks[blockIdx.x];
  const ProbDesc pd = probs[t.prob];
  const DevProgram& P = progs[t.prob];
  const int r0 = t.ti * kTile, c0 = t.tj * kTile;
  load_points<DP>(xar + pd.xoff, pd.npad, r0, xr);
  load_points<DP>(xar + pd.xoff, pd.npad, c0, xc);
  __syncthreads();
  const int i = threadIdx.x & 127, jh = threadIdx.x >> 7;
  const int gi = r0 + i;
  double xi[DP];
#pragma unroll
  for (int q = 0; q < DP; ++q) xi[q] = xr[q][i];
  bool bad = false;
  T* A = Aar + pd.moff;
  for (int jj = 0; jj < 64; ++jj) {
    const int j = jh + 2 * jj;
    const int gj = c0 + j;
    if (gi < gj) continue;
    double val;
    if (gi >= pd.n || gj >= pd.n) {
      val = gi == gj ? 1.0 : 0.0;
    } else {
      double dx[DP];
#pragma unroll
      for (int q = 0; q < DP; ++q) dx[q] = xi[q] - xc[q][j];
      Leaves L;
      val = eval_entry<DP, false>(P, dx, L);
      if (!isfinite(val)) bad = true;
      if (gi == gj) val += P.noise;
    }
    A[gi + (long)gj * pd.npad] = (T)val;
    if (DUAL) K64[pd.moff + gi + (long)gj * pd.npad] = val;
  }
  if (bad) atomicOr(info + t.prob, 2);
}

// Fused gradient contraction over every problem's lower tiles; W_ij =
// alpha_i alpha_j - Ainv_ij.  Partials per task (tasks of a problem are
// consecutive).
template <int DP, typename T, int NS>
__global__ __launch_bounds__(256, contract_occupancy(DP, NS)) void k_contract_tasks(const FillTask* __restrict__ tasks,
                                                        const ProbDesc* __restrict__ probs,
                                                        const DevProgram* __restrict__ progs,
                                                        const double* __restrict__ xar,
                                                        const T* __restrict__ Ainv,
                                                        const double* __restrict__ alpha,
                                                        double* __restrict__ partial) {
  using R = T;  // f32 matrices: entries evaluated in f32
  __shared__ double xr[DP][kTile];
  __shared__ double xc[DP][kTile];
  __shared__ double ar[kTile], ac[kTile];
  __shared__ double red[4 * kMaxAcc];
  const FillTask t = tasks[blockIdx.x];
  const ProbDesc pd = probs[t.prob];
  const DevProgram& P = progs[t.prob];
  const int r0 = t.ti * kTile, c0 = t.tj * kTile;
  load_points<DP>(xar + pd.xoff, pd.npad, r0, xr);
  load_points<DP>(xar + pd.xoff, pd.npad, c0, xc);
  if (threadIdx.x < kTile) {
    ar[threadIdx.x] = alpha[pd.voff + r0 + threadIdx.x];
    ac[threadIdx.x] = alpha[pd.voff + c0 + threadIdx.x];
  }
  __syncthreads();
  const int i = threadIdx.x & 127, jh = threadIdx.x >> 7;
  const int gi = r0 + i;
  double xi[DP];
#pragma unroll
  for (int q = 0; q < DP; ++q) xi[q] = xr[q][i];
  // per-thread sums in the matrices' type (f32 sums of a thread's at most
  // 64 f32 entries, widened to f64 at the block reduction)
  using A = T;
  A acc[NS][kAccPerLeaf];
  A accd[DP];
#pragma unroll
  for (int l = 0; l < NS; ++l) acc[l][0] = acc[l][1] = 0;
#pragma unroll
  for (int q = 0; q < DP; ++q) accd[q] = 0;
  const T* M = Ainv + pd.moff;
  if (gi < pd.n) {
    for (int jj = 0; jj < 64; ++jj) {
      const int j = jh + 2 * jj;
      const int gj = c0 + j;
      if (gj > gi || gj >= pd.n) continue;
      double w = ar[i] * ac[j] - (double)M[gi + (long)gj * pd.npad];
      if (gi != gj) w *= 2.0;
      R dx[DP];
#pragma unroll
      for (int q = 0; q < DP; ++q) dx[q] = (R)(xi[q] - xc[q][j]);
      LeavesT<R, NS> L;
      eval_entry<DP, true>(P, dx, L);
      accumulate_entry<DP>(P, dx, L, (A)w, acc, accd);
    }
  }
  store_partials<DP>(acc, accd, red, partial + (long)blockIdx.x * kMaxAcc);
}

// Sums each problem's partial rows [beg[p], end[p]) into out[p].
__global__ __launch_bounds__(256) void k_reduce_by_problem(const double* partial, const int* beg,
                                                           const int* end, double* out) {
  __shared__ double red[4];
  const int p = blockIdx.x / kMaxAcc, k = blockIdx.x % kMaxAcc;
  double acc = 0.0;
  for (int t = beg[p] + threadIdx.x; t < end[p]; t += 256) acc += partial[(long)t * kMaxAcc + k];
  acc = block_sum<256>(acc, red);
  if (threadIdx.x == 0) out[(long)p * kMaxAcc + k] = acc;
}

// Matrix-vector products over the lower 128-tiles (prob, ti, tj) of every
// problem (k_tile_mv_lower's tile body per task): task partials go to slot
// fbeg[prob] + ti (ti + 1) / 2 + tj of part (4 * 128 doubles each), which
// k_tile_mv_reduce_tasks adds per output block in a fixed order.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_tile_mv_tasks(const FillTask* __restrict__ tasks,
                                                       const ProbDesc* __restrict__ probs,
                                                       const int* __restrict__ fbeg,
                                                       const T* __restrict__ Mar,
                                                       const double* __restrict__ v,
                                                       double* __restrict__ part) {
  const FillTask t = tasks[blockIdx.x];
  const ProbDesc pd = probs[t.prob];
  const T* Mt = Mar + pd.moff + (long)t.ti * kTile + (long)t.tj * kTile * pd.npad;
  const long slot = fbeg[t.prob] + (long)t.ti * (t.ti + 1) / 2 + t.tj;
  tile_mv_tile<T, MODE>(Mt, pd.npad, v + pd.voff + (long)t.ti * kTile, v + pd.voff + (long)t.tj * kTile,
                        t.ti == t.tj, part + slot * 4 * kTile);
}

// out (and out2) over the vector arena: row v of the problem owning its
// 128-block; zero for a problem with active[p] == 0 (active may be null).
__global__ void k_tile_mv_reduce_tasks(const double* __restrict__ part,
                                       const ProbDesc* __restrict__ probs,
                                       const int* __restrict__ blk_prob, const int* __restrict__ fbeg,
                                       const int* __restrict__ active, long vtot, int mode,
                                       double* __restrict__ out, double* __restrict__ out2) {
  const long v = (long)blockIdx.x * 256 + threadIdx.x;
  if (v >= vtot) return;
  const int p = blk_prob[v / kTile];
  const ProbDesc pd = probs[p];
  double acc = 0.0, acc2 = 0.0;
  if (!active || active[p]) {
    const long local = v - pd.voff;
    tile_mv_sum(part + (long)fbeg[p] * 4 * kTile, pd.npad / kTile, (int)(local / kTile),
                (int)(local % kTile), mode, out2 != nullptr, acc, acc2);
  }
  out[v] = acc;
  if (out2) out2[v] = acc2;
}

// Per-problem scalars, one block each:
// out[p*8 + 0] = y'alpha, 1 = sum log L_ii, 2 = sum (alpha^2 - Ainv_ii) (if Ainv).
template <typename T>
__global__ __launch_bounds__(256) void k_problem_scalars(const ProbDesc* probs, const double* y,
                                                         const double* alpha, const double* logd,
                                                         const T* Ainv, double* out) {
  __shared__ double r0[4], r1[4], r2[4];
  const ProbDesc pd = probs[blockIdx.x];
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < pd.n; i += 256) {
    const double al = alpha[pd.voff + i];
    a += y[pd.voff + i] * al;
    b += logd[pd.voff + i];
    if (Ainv) c += al * al - (double)Ainv[pd.moff + (long)i * (pd.npad + 1)];
  }
  a = block_sum<256>(a, r0);
  b = block_sum<256>(b, r1);
  c = block_sum<256>(c, r2);
  if (threadIdx.x == 0) {
    out[blockIdx.x * 8 + 0] = a;
    out[blockIdx.x * 8 + 1] = b;
    if (Ainv) out[blockIdx.x * 8 + 2] = c;
  }
}

// Refinement: r = y - (p1 + p2); out[p*8 + 4..7] = max|r|, max|alpha|,
// max|y|, max row-abs-sum (q1 + q2, first pass).
__global__ __launch_bounds__(256) void k_problem_residual(const ProbDesc* probs, const double* y,
                                                          const double* p1, const double* p2,
                                                          const double* alpha, const double* q1,
                                                          const double* q2, double* r,
                                                          double* out) {
  __shared__ double m0[4], m1[4], m2[4], m3[4];
  const ProbDesc pd = probs[blockIdx.x];
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (int i = threadIdx.x; i < pd.npad; i += 256) {
    const long k = pd.voff + i;
    if (i >= pd.n) {
      r[k] = 0.0;
      continue;
    }
    const double ri = y[k] - (p2 ? p1[k] + p2[k] : p1[k]);
    r[k] = ri;
    a0 = fmax(a0, fabs(ri));
    a1 = fmax(a1, fabs(alpha[k]));
    a2 = fmax(a2, fabs(y[k]));
    if (q1) a3 = fmax(a3, q2 ? q1[k] + q2[k] : q1[k]);
  }
  a0 = block_max<256>(a0, m0);
  a1 = block_max<256>(a1, m1);
  a2 = block_max<256>(a2, m2);
  a3 = block_max<256>(a3, m3);
  if (threadIdx.x == 0) {
    out[blockIdx.x * 8 + 4] = a0;
    out[blockIdx.x * 8 + 5] = a1;
    out[blockIdx.x * 8 + 6] = a2;
    if (q1) out[blockIdx.x * 8 + 7] = a3;
  }
}

// Residual pass `it` of the captured mixed evaluation: hist[p][it] holds
// ||r||_inf and ||A||_inf ||alpha||_inf + ||y||_inf (the denominator of the
// stop test of src/precision/refine.rs:56-98) from the scalars
// k_problem_residual wrote.
__global__ void k_store_residual(const double* __restrict__ sc, double* __restrict__ hist, int K,
                                 int it, int passes) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= K) return;
  const double* q = sc + 8 * p;
  hist[(long)p * 2 * passes + 2 * it] = q[4];
  hist[(long)p * 2 * passes + 2 * it + 1] = q[7] * q[5] + q[6];
}

// alpha += d on the problems whose flag is set.
__global__ void k_masked_axpy(const ProbDesc* probs, const int* active, double* alpha,
                              const double* d) {
  const ProbDesc pd = probs[blockIdx.y];
  if (!active[blockIdx.y]) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < pd.n) alpha[pd.voff + i] += d[pd.voff + i];
}

template <typename F>
int guarded(gprx_error* err, F&& f) {
  try {
    return f();
  } catch (const DeviceError& e) {
    set_error(err, GPRX_DEVICE_ERROR, e.msg);
    return GPRX_DEVICE_ERROR;
  } catch (...) {
    set_error(err, GPRX_DEVICE_ERROR, "unexpected failure");
    return GPRX_DEVICE_ERROR;
  }
}

struct HostProb {
  HostProgram hp;
  int n, d, npad, nt;
  long xoff, voff, moff;
  std::vector<double> theta;
  // Mixed batches: the problem as given, for an FP64 re-evaluation.
  std::vector<double> x, y;
  std::vector<int32_t> ops;
};

template <typename V>
void upload(DevBuf<V>& dst, const std::vector<V>& src, hipStream_t s) {
  dst.alloc(std::max<size_t>(src.size(), 1));
  if (!src.empty())
    GPRX_HIP_CHECK(hipMemcpyAsync(dst.p, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice, s));
}

struct BatchModel : gprx_hip_model {
  BatchModel() : gprx_hip_model(kBatch) {}
  ~BatchModel() override {
    // nothing of this batch is in flight once its stream has drained
    if (stream) (void)hipStreamSynchronize(stream);
    for (gprx_hip_model* m : fb) gprx_hip_destroy(m);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (graph_b_exec) (void)hipGraphExecDestroy(graph_b_exec);
    if (graph_b) (void)hipGraphDestroy(graph_b);
    if (stream) {
      for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
      (void)hipStreamDestroy(stream);
    }
  }
  int K = 0, dp = 1, maxnt = 0, precision = GPRX_FP64;
  long xtot = 0, vtot = 0, mtot = 0;
  std::vector<HostProb> hp;
  // Mixed batches: a problem whose f32 factor fails or whose refinement
  // stalls is evaluated alone in FP64 (fb[p], created on first need and
  // kept while their matrices stay under kFallbackKeepBytes).
  static constexpr double kFallbackKeepBytes = 16.0 * (1 << 30);
  std::vector<gprx_hip_model*> fb;
  std::vector<int> fb_reason;
  double fb_bytes = 0.0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  DevBuf<double> xar, yar, logd, alpha, tv, r, dlt, p1, p2, q1, q2, K64, A64, P64, partial, accs, scal;
  DevBuf<float> A32, P32;
  DevBuf<ProbDesc> probs;
  DevBuf<DevProgram> progs;
  HostBuf<DevProgram> hprogs;
  DevBuf<int> info, active;
  HostBuf<int> hinfo, hactive;
  HostBuf<double> hscal, hacc;
  // The captured mixed evaluation refines alpha kMixedCorrections times on
  // every problem; each of its residual passes is recorded in rhist.
  static constexpr int kMixedCorrections = 2;
  DevBuf<double> rhist;
  HostBuf<double> hrhist;
  hipGraph_t graph_b = nullptr;  // mixed: X^T X, scalars, contraction
  hipGraphExec_t graph_b_exec = nullptr;
  DevBuf<FillTask> act_tasks;    // lower tiles of the problems still refining
  HostBuf<FillTask> hact;
  std::vector<FillTask> hfill;   // host copy; problem p owns [hbeg[p], hend[p])
  std::vector<int> hbeg, hend;
  // Task lists (host copy kept for the schedules).
  struct Range {
    int off = 0, count = 0;
  };
  // A launch's schedule: tile edge and its slots in d_slots.
  struct Sched {
    int tt = 128, off = 0, n = 0;
  };
  std::vector<TileTask> htasks;
  std::vector<int> task_prob;
  DevBuf<TileTask> tasks;
  // per step j of the factor: column update, panel + Y row; per distance d
  // of the inverse; X^T X
  std::vector<Range> upd, pan, inv;
  Range lauum;
  std::vector<Sched> upd_s, pan_s, inv_s;
  Sched lauum_s;
  DevBuf<int> d_slots;
  int policy = -1, small_below = -1;  // the schedule in use (-1: not built)
  bool forced = false;                // set by gprx_hip_batch_set_schedule
  double tuned_ms = 0.0;
  DevBuf<long> d_aoff, d_xoff, d_ld, d_logoff;
  DevBuf<int> d_infoidx;
  std::vector<Range> step_diag;
  DevBuf<FillTask> fill_tasks;
  int n_fill = 0;
  DevBuf<int> d_fbeg, d_fend;
  struct DimGroup {
    int dp, off, count;
    int ns;  // most leaf slots of the group's programs
  };
  std::vector<DimGroup> dp_groups;
  DevBuf<double> mvpart;        // tile matrix-vector partials, 4 * 128 per lower tile
  DevBuf<int> d_blkprob;        // problem of each 128-block of the vector arena

  bool mixed() const { return precision == GPRX_MIXED; }

  // Left-looking factor and inverse, one task per 128 x 128 output tile
  // (arena 0: K, then L; arena 1: X).  Step j of the factor:
  //   upd: A_ij -= L_i,0:j L_j,0:j^T      (i >= j; K = 128 j)
  //   diag: X_jj = L_jj^-1 and log L_jj   (k_potrf_diag_list)
  //   pan: L_ij = A_ij X_jj^T (i > j) and Y_jk = X_jj L_jk (k < j, in place:
  //        row j of L is not read again by the factor)
  // then the inverse by distance d = i - j, X_ij = -sum_{k=j}^{i-1} Y_ik X_kj
  // (L X = I with Y_ik = X_ii L_ik; K = 128 d), and A^-1 = X^T X (lower).
  void build_tasks() {
    std::vector<long> aoff, xoff, ld, logoff;
    std::vector<int> infoidx;
    htasks.clear();
    task_prob.clear();
    auto open_range = [&](Range& rg) { rg.off = (int)htasks.size(); };
    auto close_range = [&](Range& rg) { rg.count = (int)htasks.size() - rg.off; };
    auto add = [&](int p, TileTask t) {
      htasks.push_back(t);
      task_prob.push_back(p);
    };
    upd.assign(maxnt, Range{});
    pan.assign(maxnt, Range{});
    inv.assign(maxnt, Range{});
    step_diag.assign(maxnt, Range{});
    for (int j = 0; j < maxnt; ++j) {
      step_diag[j].off = (int)aoff.size();
      for (int p = 0; p < K; ++p) {
        const HostProb& h = hp[p];
        if (h.nt <= j) continue;
        const long L = h.npad, jj = (long)j * kTile;
        aoff.push_back(h.moff + jj + jj * L);
        xoff.push_back(h.moff + jj + jj * L);
        ld.push_back(L);
        logoff.push_back(h.voff + jj);
        infoidx.push_back(p);
      }
      step_diag[j].count = (int)aoff.size() - step_diag[j].off;
      open_range(upd[j]);
      if (j > 0)
        for (int p = 0; p < K; ++p) {
          const HostProb& h = hp[p];
          if (h.nt <= j) continue;
          const long L = h.npad;
          auto at = [&](int r, int c) { return h.moff + (long)r * kTile + (long)c * kTile * L; };
          for (int i = j; i < h.nt; ++i)
            add(p, TileTask{at(i, 0), at(j, 0), at(i, j), (int)L, j * kTile, 1, 0, i == j, 0, -1.f, 1.f});
        }
      close_range(upd[j]);
      open_range(pan[j]);
      for (int p = 0; p < K; ++p) {
        const HostProb& h = hp[p];
        if (h.nt <= j) continue;
        const long L = h.npad;
        auto at = [&](int r, int c) { return h.moff + (long)r * kTile + (long)c * kTile * L; };
        for (int i = j + 1; i < h.nt; ++i)  // L_ij = A_ij X_jj^T: B from arena 1
          add(p, TileTask{at(i, j), at(j, j), at(i, j), (int)L, kTile, 1, 0, 0, 2, 1.f, 0.f});
        for (int k = 0; k < j; ++k)  // Y_jk = X_jj L_jk: A from arena 1
          add(p, TileTask{at(j, j), at(j, k), at(j, k), (int)L, kTile, 0, 0, 0, 1, 1.f, 0.f});
      }
      close_range(pan[j]);
    }
    for (int d = 1; d < maxnt; ++d) {
      open_range(inv[d]);
      for (int p = 0; p < K; ++p) {
        const HostProb& h = hp[p];
        const long L = h.npad;
        auto at = [&](int r, int c) { return h.moff + (long)r * kTile + (long)c * kTile * L; };
        for (int j = 0; j + d < h.nt; ++j)  // Y from arena 0, X (B and C) in arena 1
          add(p, TileTask{at(j + d, j), at(j, j), at(j + d, j), (int)L, d * kTile, 0, 0, 0, 6,
                          -1.f, 0.f});
      }
      close_range(inv[d]);
    }
    // A^-1 (lower) = X^T X: C_ij = sum_{k >= i} X_ki^T X_kj, into arena 0.
    // A problem's tiles share the end of their K-range (its last row), so
    // they run from it down: tiles started together read the same rows of X
    // at the same time (as the single-problem X^T X does).
    open_range(lauum);
    for (int p = 0; p < K; ++p) {
      const HostProb& h = hp[p];
      const long L = h.npad;
      for (int i = 0; i < h.nt; ++i)
        for (int j = 0; j <= i; ++j) {
          const long a = h.moff + (long)i * kTile + (long)i * kTile * L;
          const long b = h.moff + (long)i * kTile + (long)j * kTile * L;
          const long c = h.moff + (long)i * kTile + (long)j * kTile * L;
          add(p, TileTask{a, b, c, (int)L, (h.nt - i) * kTile, 0, 1, (i == j ? 1 : 0) | 2, 3, 1.f, 0.f});
        }
    }
    close_range(lauum);
    upload(tasks, htasks, stream);
    upload(d_aoff, aoff, stream);
    upload(d_xoff, xoff, stream);
    upload(d_ld, ld, stream);
    upload(d_logoff, logoff, stream);
    upload(d_infoidx, infoidx, stream);
    // Fill / contraction tiles: problems grouped by padded dimension (one
    // launch per group, so a d = 2 problem does not pay for d = 16), each
    // problem's tiles contiguous.
    std::vector<int> order(K);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return padded_dim(hp[a].d) < padded_dim(hp[b].d); });
    std::vector<FillTask> ft;
    std::vector<int> fbeg(K, 0), fend(K, 0);
    dp_groups.clear();
    for (int p : order) {
      const int pd = padded_dim(hp[p].d);
      if (dp_groups.empty() || dp_groups.back().dp != pd) dp_groups.push_back(DimGroup{pd, (int)ft.size(), 0, 0});
      dp_groups.back().ns = std::max(dp_groups.back().ns, hp[p].hp.ns);
      fbeg[p] = (int)ft.size();
      for (int i = 0; i < hp[p].nt; ++i)
        for (int j = 0; j <= i; ++j) ft.push_back(FillTask{p, i, j});
      fend[p] = (int)ft.size();
      dp_groups.back().count = (int)ft.size() - dp_groups.back().off;
    }
    n_fill = (int)ft.size();
    upload(fill_tasks, ft, stream);
    upload(d_fbeg, fbeg, stream);
    upload(d_fend, fend, stream);
    hfill = ft;
    hbeg = fbeg;
    hend = fend;
  }

  // Slots of one launch under a schedule policy.  Items are tasks (tile
  // edge 128) or their quadrants (64; the upper quadrant of a diagonal tile
  // is skipped); an item's work is its K in 128-blocks plus one.
  //   policy 0: build order (problem-major), one contiguous run per XCD
  //   policy 1: longest first, each to the least-loaded XCD
  //   policy 2: problems as units, heaviest first to the least-loaded XCD,
  //             a problem's items longest first (its panels share an L2)
  // Slot 8k + x holds XCD x's k-th item; -1 pads.
  std::vector<int> build_slots(const Range& rg, int tt, int pol) const {
    struct Item {
      int code, prob;
      long work;
    };
    std::vector<Item> items;
    for (int t = rg.off; t < rg.off + rg.count; ++t) {
      const TileTask& tk = htasks[t];
      const long w = tk.K / kTile + 1;
      if (tt == 128) {
        items.push_back(Item{t << 2, task_prob[t], w});
      } else {
        for (int q = 0; q < 4; ++q) {
          if ((tk.lower & 1) && q == 2) continue;  // rows 0-63, columns 64-127: above the diagonal
          items.push_back(Item{t << 2 | q, task_prob[t], w});
        }
      }
    }
    constexpr int kX = 8;
    std::vector<std::vector<int>> lists(kX);
    const int total = (int)items.size();
    if (pol == 0) {
      const int q = total / kX, r = total % kX;
      int pos = 0;
      for (int x = 0; x < kX; ++x) {
        const int len = q + (x < r ? 1 : 0);
        for (int k = 0; k < len; ++k) lists[x].push_back(items[pos + k].code);
        pos += len;
      }
    } else if (pol == 1) {
      std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.work > b.work; });
      long load[kX] = {0};
      for (const Item& it : items) {
        int x = 0;
        for (int i = 1; i < kX; ++i)
          if (load[i] < load[x]) x = i;
        load[x] += it.work;
        lists[x].push_back(it.code);
      }
    } else {
      std::vector<std::vector<Item>> groups(K);
      for (const Item& it : items) groups[it.prob].push_back(it);
      std::vector<int> order;
      std::vector<long> gw(K, 0);
      for (int p = 0; p < K; ++p) {
        for (const Item& it : groups[p]) gw[p] += it.work;
        if (!groups[p].empty()) order.push_back(p);
        std::stable_sort(groups[p].begin(), groups[p].end(),
                         [](const Item& a, const Item& b) { return a.work > b.work; });
      }
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return gw[a] > gw[b]; });
      long load[kX] = {0};
      for (int p : order) {
        int x = 0;
        for (int i = 1; i < kX; ++i)
          if (load[i] < load[x]) x = i;
        load[x] += gw[p];
        for (const Item& it : groups[p]) lists[x].push_back(it.code);
      }
    }
    size_t len = 0;
    for (auto& l : lists) len = std::max(len, l.size());
    std::vector<int> out(len * kX, -1);
    for (int x = 0; x < kX; ++x)
      for (size_t k = 0; k < lists[x].size(); ++k) out[k * kX + x] = lists[x][k];
    return out;
  }

  // Builds and uploads the slots of every launch for (policy, small_below):
  // launches with fewer than small_below tasks run 64 x 64 sub-tiles, except
  // the panel launches, whose tasks overwrite the tile they read (a quadrant
  // would clobber rows its sibling quadrants still read).
  void set_schedule(int pol, int below) {
    std::vector<int> all;
    auto one = [&](const Range& rg, Sched& sc, bool in_place) {
      sc.tt = (!in_place && rg.count < below) ? 64 : 128;
      sc.off = (int)all.size();
      const std::vector<int> sl = rg.count ? build_slots(rg, sc.tt, pol) : std::vector<int>{};
      sc.n = (int)sl.size();
      all.insert(all.end(), sl.begin(), sl.end());
    };
    upd_s.assign(maxnt, Sched{});
    pan_s.assign(maxnt, Sched{});
    inv_s.assign(maxnt, Sched{});
    for (int j = 0; j < maxnt; ++j) {
      one(upd[j], upd_s[j], false);
      one(pan[j], pan_s[j], true);
      one(inv[j], inv_s[j], false);
    }
    one(lauum, lauum_s, false);
    if (all.empty()) all.push_back(-1);
    if (d_slots.n < all.size()) {
      d_slots.release();
      d_slots.alloc(all.size());
    }
    GPRX_HIP_CHECK(hipMemcpyAsync(d_slots.p, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice, stream));
    GPRX_HIP_CHECK(hipStreamSynchronize(stream));
    policy = pol;
    small_below = below;
  }

  template <typename T>
  void launch_tasks(const Sched& sc, T* X0, T* X1) {
    if (sc.n == 0) return;
    if (sc.tt == 64)
      hipLaunchKernelGGL((k_gemm_tasks<T, 64>), dim3(sc.n), dim3(256), 0, stream, tasks.p,
                         d_slots.p + sc.off, X0, X1);
    else
      hipLaunchKernelGGL((k_gemm_tasks<T, 128>), dim3(sc.n), dim3(256), 0, stream, tasks.p,
                         d_slots.p + sc.off, X0, X1);
    GPRX_HIP_CHECK(hipGetLastError());
  }

  // out (out2) = a tile product (kTmvRow / kTmvCol / kTmvSym / kTmvSymAbs)
  // over the lower tiles of every problem, or over the n_act tiles in
  // act_tasks, with out zero for the other problems.
  template <typename T>
  void tile_mv(const T* M, const double* v, double* out, int mode, double* out2 = nullptr, int n_act = -1) {
    const int count = n_act < 0 ? n_fill : n_act;
    const FillTask* t = n_act < 0 ? fill_tasks.p : act_tasks.p;
    if (count > 0) {
      switch (mode) {
        case kTmvRow:
          hipLaunchKernelGGL((k_tile_mv_tasks<T, kTmvRow>), dim3(count), dim3(256), 0, stream, t, probs.p,
                             d_fbeg.p, M, v, mvpart.p);
          break;
        case kTmvCol:
          hipLaunchKernelGGL((k_tile_mv_tasks<T, kTmvCol>), dim3(count), dim3(256), 0, stream, t, probs.p,
                             d_fbeg.p, M, v, mvpart.p);
          break;
        case kTmvSym:
          hipLaunchKernelGGL((k_tile_mv_tasks<T, kTmvSym>), dim3(count), dim3(256), 0, stream, t, probs.p,
                             d_fbeg.p, M, v, mvpart.p);
          break;
        default:
          hipLaunchKernelGGL((k_tile_mv_tasks<T, kTmvSymAbs>), dim3(count), dim3(256), 0, stream, t,
                             probs.p, d_fbeg.p, M, v, mvpart.p);
          break;
      }
    }
    hipLaunchKernelGGL(k_tile_mv_reduce_tasks, dim3((unsigned)((vtot + 255) / 256)), dim3(256), 0, stream,
                       mvpart.p, probs.p, d_blkprob.p, d_fbeg.p, n_act < 0 ? nullptr : active.p, vtot,
                       mode & 3, out, mode == kTmvSymAbs ? out2 : nullptr);
    GPRX_HIP_CHECK(hipGetLastError());
  }

  // Uploads the lower tiles of the problems with act[p] set; returns their
  // count.  The pinned staging buffer is rewritten only after the stream has
  // been synchronised (the refinement loop reads its residuals each pass).
  // Problem p alone in the FP64 form (default jitter policy fixed(0), as the
  // batch's own factor): the mixed batch's fallback.
  int eval_fp64_alone(int p, const double* th, double* nlml, double* grad, gprx_error* e) {
    HostProb& h = hp[p];
    const size_t np = h.theta.size();
    if (!fb[p]) {
      const gprx_kernel_program prog{h.ops.data(), h.ops.size() / 2};
      const int rc = gprx_hip_exact_create(h.x.data(), (size_t)h.n, (size_t)h.d, h.y.data(), &prog, th,
                                           np, GPRX_FP64, nullptr, &fb[p], e);
      if (rc != GPRX_OK) return rc;
    }
    const int rc = gprx_hip_exact_value_and_grad(fb[p], th, np, nlml, grad, e);
    const double bytes = 3.0 * sizeof(double) * (double)h.npad * (double)h.npad;
    if (fb_bytes + bytes <= kFallbackKeepBytes) {
      fb_bytes += bytes;
    } else {
      gprx_hip_destroy(fb[p]);
      fb[p] = nullptr;
    }
    return rc;
  }

  // Uploads the lower tiles of the problems with act[p] set; returns their
  // count (the restricted matrix-vector products of the refinement).
  int set_active(const std::vector<int>& act) {
    int c = 0;
    hact.alloc(std::max(n_fill, 1));
    for (int p = 0; p < K; ++p)
      if (act[p])
        for (int t = hbeg[p]; t < hend[p]; ++t) hact.p[c++] = hfill[t];
    if (c > 0)
      GPRX_HIP_CHECK(hipMemcpyAsync(act_tasks.p, hact.p, sizeof(FillTask) * c, hipMemcpyHostToDevice, stream));
    return c;
  }

  template <typename T>
  void fill(T* Aar) {
    const bool dual = mixed();
    for (const DimGroup& g : dp_groups) {
#define GPRX_FILL(DPV, DUAL)                                                                        \
  hipLaunchKernelGGL((k_fill_tasks<DPV, T, DUAL>), dim3(g.count), dim3(256), 0, stream,             \
                     fill_tasks.p + g.off, probs.p, progs.p, xar.p, Aar, K64.p, info.p)
      switch (g.dp) {
        case 1: if (dual) GPRX_FILL(1, true); else GPRX_FILL(1, false); break;
        case 2: if (dual) GPRX_FILL(2, true); else GPRX_FILL(2, false); break;
        case 4: if (dual) GPRX_FILL(4, true); else GPRX_FILL(4, false); break;
        case 8: if (dual) GPRX_FILL(8, true); else GPRX_FILL(8, false); break;
        case 16: if (dual) GPRX_FILL(16, true); else GPRX_FILL(16, false); break;
        default: if (dual) GPRX_FILL(32, true); else GPRX_FILL(32, false); break;
      }
#undef GPRX_FILL
    }
    GPRX_HIP_CHECK(hipGetLastError());
  }

  template <typename T>
  void contract(const T* Ainv) {
    for (const DimGroup& g : dp_groups) {
#define GPRX_CON_NS(DPV, NSV)                                                                      \
  hipLaunchKernelGGL((k_contract_tasks<DPV, T, NSV>), dim3(g.count), dim3(256), 0, stream,          \
                     fill_tasks.p + g.off, probs.p, progs.p, xar.p, Ainv, alpha.p,                 \
                     partial.p + (long)g.off * kMaxAcc)
#define GPRX_CON(DPV)                                                                              \
  if (g.ns <= 2)                                                                                   \
    GPRX_CON_NS(DPV, 2);                                                                           \
  else if (g.ns <= 4)                                                                              \
    GPRX_CON_NS(DPV, 4);                                                                           \
  else                                                                                             \
    GPRX_CON_NS(DPV, kMaxSlots)
      switch (g.dp) {
        case 1: GPRX_CON(1); break;
        case 2: GPRX_CON(2); break;
        case 4: GPRX_CON(4); break;
        case 8: GPRX_CON(8); break;
        case 16: GPRX_CON(16); break;
        default: GPRX_CON(32); break;
      }
#undef GPRX_CON
#undef GPRX_CON_NS
    }
    hipLaunchKernelGGL(k_reduce_by_problem, dim3(K * kMaxAcc), dim3(256), 0, stream, partial.p,
                       d_fbeg.p, d_fend.p, accs.p);
    GPRX_HIP_CHECK(hipGetLastError());
  }

  // Factor + inverse of every problem: L (then Y) in Aar, X in Par.
  template <typename T>
  void factor(T* Aar, T* Par) {
    for (int j = 0; j < maxnt; ++j) {
      launch_tasks<T>(upd_s[j], Aar, Par);
      const Range& dg = step_diag[j];
      potrf_diag_list<T>(Aar, Par, d_aoff.p + dg.off, d_xoff.p + dg.off, d_ld.p + dg.off,
                         d_logoff.p + dg.off, logd.p, info.p, d_infoidx.p + dg.off, dg.count, stream);
      launch_tasks<T>(pan_s[j], Aar, Par);
    }
    for (int d = 1; d < maxnt; ++d) launch_tasks<T>(inv_s[d], Aar, Par);
  }

  // The schedule: timed once per candidate on this batch's own factor and
  // X^T X (the launches the schedule shapes), the fastest kept.  Candidates:
  // the three policies of build_slots, each with 64-tiles below 0, 256 and
  // 1024 tasks per launch.  A pick is shared by batches of the same shape
  // in the process.
  void tune() {
    if (policy >= 0) return;
    const std::string key = shape_key();
    {
      std::lock_guard<std::mutex> lock(tune_mu());
      auto it = tune_cache().find(key);
      if (it != tune_cache().end()) {
        set_schedule(it->second.first, it->second.second);
        return;
      }
    }
    hipStream_t s = stream;
    GPRX_HIP_CHECK(hipMemcpyAsync(progs.p, hprogs.p, sizeof(DevProgram) * K, hipMemcpyHostToDevice, s));
    double best = 1e300;
    int bp = 0, bb = 0;
    for (int pol = 0; pol < 3; ++pol)
      for (int below : {0, 256, 1024}) {
        set_schedule(pol, below);
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // first pass warms caches and code
          if (mixed()) {
            fill<float>(A32.p);
            GPRX_HIP_CHECK(hipEventRecord(ev[0], s));
            factor<float>(A32.p, P32.p);
            launch_tasks<float>(lauum_s, A32.p, P32.p);
          } else {
            fill<double>(A64.p);
            GPRX_HIP_CHECK(hipEventRecord(ev[0], s));
            factor<double>(A64.p, P64.p);
            launch_tasks<double>(lauum_s, A64.p, P64.p);
          }
          GPRX_HIP_CHECK(hipEventRecord(ev[1], s));
          GPRX_HIP_CHECK(hipEventSynchronize(ev[1]));
          GPRX_HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
        }
        if (ms < best) {
          best = ms;
          bp = pol;
          bb = below;
        }
      }
    set_schedule(bp, bb);
    tuned_ms = best;
    std::lock_guard<std::mutex> lock(tune_mu());
    tune_cache()[key] = {bp, bb};
  }

  std::string shape_key() const {
    std::string k = std::to_string(precision) + ":" + std::to_string(dp);
    for (const HostProb& h : hp) k += "," + std::to_string(h.nt);
    return k;
  }
  static std::mutex& tune_mu() {
    static std::mutex m;
    return m;
  }
  static std::map<std::string, std::pair<int, int>>& tune_cache() {
    static std::map<std::string, std::pair<int, int>> c;
    return c;
  }

  // The device part of one FP64 evaluation (capturable: kernels and
  // device copies only).
  void record_fp64() {
    fill<double>(A64.p);
    factor<double>(A64.p, P64.p);
    tile_mv<double>(P64.p, yar.p, tv.p, kTmvRow);     // X y
    tile_mv<double>(P64.p, tv.p, alpha.p, kTmvCol);   // X^T (X y)
    launch_tasks<double>(lauum_s, A64.p, P64.p);
    hipLaunchKernelGGL((k_problem_scalars<double>), dim3(K), dim3(256), 0, stream, probs.p, yar.p,
                       alpha.p, logd.p, A64.p, scal.p);
    contract<double>(A64.p);
  }

  // The device parts of one mixed evaluation (capturable).  Part A: f32
  // factor and inverse, alpha through the f32 inverse, then
  // kMixedCorrections corrections with f64 residuals on every problem, the
  // residual of each pass recorded (kMixedCorrections + 1 passes; r holds
  // the last).  Part B, after the refinement has finished: X^T X, the
  // scalars and the contraction.  `active` holds 1 for every problem.
  void record_mixed_a() {
    fill<float>(A32.p);
    factor<float>(A32.p, P32.p);
    tile_mv<float>(P32.p, yar.p, tv.p, kTmvRow);
    tile_mv<float>(P32.p, tv.p, alpha.p, kTmvCol);
    const int passes = kMixedCorrections + 1;
    for (int it = 0; it < passes; ++it) {
      // p1 = A alpha over the lower triangle; on the first pass also the
      // absolute row sums (max: ||A||_inf)
      tile_mv<double>(K64.p, alpha.p, p1.p, it == 0 ? kTmvSymAbs : kTmvSym, q1.p);
      hipLaunchKernelGGL(k_problem_residual, dim3(K), dim3(256), 0, stream, probs.p, yar.p, p1.p,
                         nullptr, alpha.p, it == 0 ? q1.p : nullptr, nullptr, r.p, scal.p);
      hipLaunchKernelGGL(k_store_residual, dim3((K + 255) / 256), dim3(256), 0, stream, scal.p,
                         rhist.p, K, it, passes);
      if (it == kMixedCorrections) break;
      correct_alpha(-1);
    }
    GPRX_HIP_CHECK(hipGetLastError());
  }
  void record_mixed_b() {
    launch_tasks<float>(lauum_s, A32.p, P32.p);
    hipLaunchKernelGGL((k_problem_scalars<float>), dim3(K), dim3(256), 0, stream, probs.p, yar.p,
                       alpha.p, logd.p, A32.p, scal.p);
    contract<float>(A32.p);
  }
  // alpha += X^T (X r) on the problems flagged in `active` (n_act tiles in
  // act_tasks, or every problem when n_act < 0)
  void correct_alpha(int n_act) {
    tile_mv<float>(P32.p, r.p, tv.p, kTmvRow, nullptr, n_act);
    tile_mv<float>(P32.p, tv.p, dlt.p, kTmvCol, nullptr, n_act);
    hipLaunchKernelGGL(k_masked_axpy, dim3((65536 + 255) / 256, K), dim3(256), 0, stream, probs.p,
                       active.p, alpha.p, dlt.p);
    GPRX_HIP_CHECK(hipGetLastError());
  }
};

}  // namespace

extern "C" {

int gprx_hip_batch_create(const gprx_problem* problems, size_t k, int32_t precision,
                          gprx_hip_model** out, gprx_error* err) {
  if (out) *out = nullptr;
  return guarded(err, [&]() -> int {
    if (!problems || !out || k == 0) {
      set_error(err, GPRX_EMPTY_INPUT, "no problems");
      return GPRX_EMPTY_INPUT;
    }
    if (precision != GPRX_FP64 && precision != GPRX_MIXED) {
      set_error(err, GPRX_UNSUPPORTED, "unknown precision");
      return GPRX_UNSUPPORTED;
    }
    auto mo = std::make_unique<BatchModel>();
    BatchModel& B = *mo;
    B.K = (int)k;
    B.precision = precision;
    B.hp.resize(k);
    B.fb.assign(k, nullptr);
    B.fb_reason.assign(k, GPRX_FALLBACK_NONE);
    int maxd = 1;
    for (size_t p = 0; p < k; ++p) {
      const gprx_problem& pr = problems[p];
      if (!pr.x || !pr.y || pr.n == 0 || pr.d == 0) {
        set_error(err, GPRX_EMPTY_INPUT, "empty problem " + std::to_string(p));
        return GPRX_EMPTY_INPUT;
      }
      if (pr.d > (size_t)kMaxDim || pr.n > 65536) {
        set_error(err, GPRX_UNSUPPORTED, "batched problems take d <= 32 and n <= 65536");
        return GPRX_UNSUPPORTED;
      }
      if (!all_finite(pr.x, pr.n * pr.d) || !all_finite(pr.y, pr.n)) {
        set_error(err, GPRX_NON_FINITE_INPUT, "non-finite data in problem " + std::to_string(p));
        return GPRX_NON_FINITE_INPUT;
      }
      HostProb& h = B.hp[p];
      std::string why;
      int rc = parse_program(pr.kernel, (int)pr.d, &h.hp, &why);
      if (rc != GPRX_OK) {
        set_error(err, rc, why + " (problem " + std::to_string(p) + ")");
        return rc;
      }
      if (!pr.log_theta) {
        set_error(err, GPRX_LENGTH_MISMATCH, "missing parameters");
        return GPRX_LENGTH_MISMATCH;
      }
      rc = exact_validate_theta(h.hp, pr.log_theta, pr.p, err);
      if (rc != GPRX_OK) return rc;
      h.n = (int)pr.n;
      h.d = (int)pr.d;
      h.npad = (int)round_up(pr.n, kTile);
      h.nt = h.npad / kTile;
      h.theta.assign(pr.log_theta, pr.log_theta + pr.p);
      if (precision == GPRX_MIXED) {
        h.x.assign(pr.x, pr.x + pr.n * pr.d);
        h.y.assign(pr.y, pr.y + pr.n);
        h.ops.assign(pr.kernel->ops, pr.kernel->ops + 2 * pr.kernel->n_ops);
      }
      maxd = std::max(maxd, h.d);
    }
    B.dp = padded_dim(maxd);
    for (auto& h : B.hp) {
      h.xoff = B.xtot;
      h.voff = B.vtot;
      h.moff = B.mtot;
      B.xtot += (long)h.npad * B.dp;
      B.vtot += h.npad;
      B.mtot += (long)h.npad * h.npad;
      B.maxnt = std::max(B.maxnt, h.nt);
    }
    GPRX_HIP_CHECK(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
    for (auto& e : B.ev) GPRX_HIP_CHECK(hipEventCreate(&e));
    hipStream_t s = B.stream;
    {
      std::vector<double> xh(B.xtot, 0.0), yh(B.vtot, 0.0);
      std::vector<ProbDesc> pdv(k);
      for (size_t p = 0; p < k; ++p) {
        const gprx_problem& pr = problems[p];
        const HostProb& h = B.hp[p];
        for (size_t q = 0; q < pr.d; ++q)
          for (size_t i = 0; i < pr.n; ++i) xh[h.xoff + q * h.npad + i] = pr.x[q * pr.n + i];
        for (size_t i = 0; i < pr.n; ++i) yh[h.voff + i] = pr.y[i];
        pdv[p] = ProbDesc{h.xoff, h.voff, h.moff, h.n, h.npad};
      }
      upload(B.xar, xh, s);
      upload(B.yar, yh, s);
      upload(B.probs, pdv, s);
      GPRX_HIP_CHECK(hipStreamSynchronize(s));
    }
    for (DevBuf<double>* b : {&B.logd, &B.alpha, &B.tv, &B.r, &B.dlt, &B.p1, &B.p2, &B.q1, &B.q2}) {
      b->alloc(B.vtot);
      GPRX_HIP_CHECK(hipMemsetAsync(b->p, 0, sizeof(double) * B.vtot, s));
    }
    if (precision == GPRX_MIXED) {
      B.A32.alloc(B.mtot);
      B.P32.alloc(B.mtot);
      B.K64.alloc(B.mtot);
      GPRX_HIP_CHECK(hipMemsetAsync(B.A32.p, 0, sizeof(float) * B.mtot, s));
      GPRX_HIP_CHECK(hipMemsetAsync(B.P32.p, 0, sizeof(float) * B.mtot, s));
      GPRX_HIP_CHECK(hipMemsetAsync(B.K64.p, 0, sizeof(double) * B.mtot, s));
    } else {
      B.A64.alloc(B.mtot);
      B.P64.alloc(B.mtot);
      GPRX_HIP_CHECK(hipMemsetAsync(B.A64.p, 0, sizeof(double) * B.mtot, s));
      GPRX_HIP_CHECK(hipMemsetAsync(B.P64.p, 0, sizeof(double) * B.mtot, s));
    }
    B.progs.alloc(k);
    B.hprogs.alloc(k);
    B.info.alloc(k);
    B.active.alloc(k);
    B.hinfo.alloc(k);
    B.hactive.alloc(k);
    B.rhist.alloc((size_t)k * 2 * (BatchModel::kMixedCorrections + 1));
    B.hrhist.alloc((size_t)k * 2 * (BatchModel::kMixedCorrections + 1));
    B.scal.alloc(8 * k);
    B.hscal.alloc(8 * k);
    B.accs.alloc(k * kMaxAcc);
    B.hacc.alloc(k * kMaxAcc);
    B.build_tasks();
    B.act_tasks.alloc(std::max(B.n_fill, 1));
    B.mvpart.alloc((size_t)std::max(B.n_fill, 1) * 4 * kTile);
    {
      std::vector<int> bp(B.vtot / kTile);
      for (int p = 0; p < B.K; ++p)
        for (long q = B.hp[p].voff / kTile; q < (B.hp[p].voff + B.hp[p].npad) / kTile; ++q) bp[q] = p;
      upload(B.d_blkprob, bp, s);
    }
    prepare_dense_kernels();
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    B.partial.alloc((size_t)B.n_fill * kMaxAcc);
    *out = mo.release();
    return GPRX_OK;
  });
}

int gprx_hip_batch_value_and_grad(gprx_hip_model* model, const double* const* log_theta,
                                  double* nlml, double* const* grad, gprx_error* errs) {
  gprx_error* err0 = errs;
  return guarded(err0, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kBatch || !log_theta || !nlml) {
      set_error(err0, GPRX_LENGTH_MISMATCH, "bad handle or null output");
      return GPRX_LENGTH_MISMATCH;
    }
    auto* B = static_cast<BatchModel*>(model);
    hipStream_t s = B->stream;
    std::vector<int> ok(B->K, 1);
    std::fill(B->fb_reason.begin(), B->fb_reason.end(), GPRX_FALLBACK_NONE);
    for (int p = 0; p < B->K; ++p) {
      HostProb& h = B->hp[p];
      gprx_error* e = errs ? errs + p : nullptr;
      if (e) std::memset(e, 0, sizeof(*e));
      if (exact_validate_theta(h.hp, log_theta[p], h.theta.size(), e) != GPRX_OK) {
        ok[p] = 0;
        make_dev_program(h.hp, h.theta.data(), std::exp(h.theta.back()), B->hprogs.p + p);
      } else {
        make_dev_program(h.hp, log_theta[p], std::exp(log_theta[p][h.hp.n_params]), B->hprogs.p + p);
      }
    }
    GPRX_HIP_CHECK(hipMemcpyAsync(B->progs.p, B->hprogs.p, sizeof(DevProgram) * B->K, hipMemcpyHostToDevice, s));
    B->tune();
    GPRX_HIP_CHECK(hipMemsetAsync(B->info.p, 0, sizeof(int) * B->K, s));
    GPRX_HIP_CHECK(hipEventRecord(B->ev[0], s));
    const bool fp64 = B->precision == GPRX_FP64;
    if (!fp64) {  // graph A's corrections apply to every problem
      for (int p = 0; p < B->K; ++p) B->hactive.p[p] = 1;
      GPRX_HIP_CHECK(hipMemcpyAsync(B->active.p, B->hactive.p, sizeof(int) * B->K, hipMemcpyHostToDevice, s));
    }
    if (!B->graph_exec) {
      // Capture one evaluation (the mixed form: its two parts); replays
      // reuse it with new parameters.
      GPRX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      if (fp64)
        B->record_fp64();
      else
        B->record_mixed_a();
      GPRX_HIP_CHECK(hipStreamEndCapture(s, &B->graph));
      GPRX_HIP_CHECK(hipGraphInstantiate(&B->graph_exec, B->graph, nullptr, nullptr, 0));
      if (!fp64) {
        GPRX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        B->record_mixed_b();
        GPRX_HIP_CHECK(hipStreamEndCapture(s, &B->graph_b));
        GPRX_HIP_CHECK(hipGraphInstantiate(&B->graph_b_exec, B->graph_b, nullptr, nullptr, 0));
      }
    }
    GPRX_HIP_CHECK(hipGraphLaunch(B->graph_exec, s));
    if (!fp64) {
      // One readback of the recorded residual passes; the stop rules of
      // src/precision/refine.rs:56-98 (converged: ||r|| < 10 n eps_64 (||A||
      // ||alpha|| + ||y||); stalled: two ratios above 0.9; at most 10
      // residual passes) applied to them, and the problems still refining
      // go on uncaptured, restricted to their own tiles.
      const int passes = BatchModel::kMixedCorrections + 1;
      GPRX_HIP_CHECK(hipMemcpyAsync(B->hrhist.p, B->rhist.p, sizeof(double) * 2 * passes * B->K,
                                    hipMemcpyDeviceToHost, s));
      GPRX_HIP_CHECK(hipMemcpyAsync(B->hinfo.p, B->info.p, sizeof(int) * B->K, hipMemcpyDeviceToHost, s));
      GPRX_HIP_CHECK(hipStreamSynchronize(s));
      std::vector<int> act(B->K, 0), conv(B->K, 0), streak(B->K, 0);
      std::vector<double> prev(B->K, -1.0);
      // one residual pass of problem p: returns false once p stops refining
      auto step = [&](int p, double rinf, double den) {
        const double tol = 10.0 * B->hp[p].n * 2.220446049250313e-16;
        if (den > 0.0 && rinf / den < tol) {
          conv[p] = 1;
          return false;
        }
        if (prev[p] >= 0.0) {
          const double ratio = prev[p] > 0.0 ? rinf / prev[p] : 0.0;
          if (ratio > 0.9) {
            if (++streak[p] >= 2) return false;
          } else {
            streak[p] = 0;
          }
        }
        prev[p] = rinf;
        return true;
      };
      int n_live = 0;
      for (int p = 0; p < B->K; ++p) {
        if (!ok[p] || B->hinfo.p[p]) continue;
        bool live = true;
        for (int it = 0; it < passes && live; ++it) {
          const double* hh = B->hrhist.p + (long)p * 2 * passes + 2 * it;
          live = step(p, hh[0], hh[1]);
        }
        act[p] = live ? 1 : 0;
        n_live += act[p];
      }
      for (int it = passes; it < 10 && n_live > 0; ++it) {
        for (int p = 0; p < B->K; ++p) B->hactive.p[p] = act[p];
        GPRX_HIP_CHECK(hipMemcpyAsync(B->active.p, B->hactive.p, sizeof(int) * B->K, hipMemcpyHostToDevice, s));
        const int n_act = B->set_active(act);
        B->correct_alpha(n_act);
        B->tile_mv<double>(B->K64.p, B->alpha.p, B->p1.p, kTmvSym, B->q1.p, n_act);
        hipLaunchKernelGGL(k_problem_residual, dim3(B->K), dim3(256), 0, s, B->probs.p, B->yar.p,
                           B->p1.p, nullptr, B->alpha.p, nullptr, nullptr, B->r.p, B->scal.p);
        GPRX_HIP_CHECK(hipMemcpyAsync(B->hscal.p, B->scal.p, sizeof(double) * 8 * B->K, hipMemcpyDeviceToHost, s));
        GPRX_HIP_CHECK(hipStreamSynchronize(s));
        n_live = 0;
        for (int p = 0; p < B->K; ++p) {
          if (!act[p]) continue;
          const double* sc = B->hscal.p + 8 * p;
          act[p] = step(p, sc[4], sc[7] * sc[5] + sc[6]) ? 1 : 0;
          n_live += act[p];
        }
      }
      for (int p = 0; p < B->K; ++p) {
        if (!ok[p]) continue;
        if (B->hinfo.p[p])
          B->fb_reason[p] = GPRX_FALLBACK_JITTER;
        else if (!conv[p])
          B->fb_reason[p] = GPRX_FALLBACK_REFINEMENT;
      }
      GPRX_HIP_CHECK(hipGraphLaunch(B->graph_b_exec, s));
    }
    GPRX_HIP_CHECK(hipEventRecord(B->ev[1], s));
    GPRX_HIP_CHECK(hipMemcpyAsync(B->hscal.p, B->scal.p, sizeof(double) * 8 * B->K, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(B->hacc.p, B->accs.p, sizeof(double) * kMaxAcc * B->K, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipMemcpyAsync(B->hinfo.p, B->info.p, sizeof(int) * B->K, hipMemcpyDeviceToHost, s));
    GPRX_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, B->ev[0], B->ev[1]);
    B->last_eval_ms = ms;
    for (int p = 0; p < B->K; ++p) {
      HostProb& h = B->hp[p];
      gprx_error* e = errs ? errs + p : nullptr;
      const int P = h.hp.n_params;
      if (!ok[p]) {
        nlml[p] = NAN;
        continue;
      }
      if (B->fb_reason[p] != GPRX_FALLBACK_NONE) {
        if (B->eval_fp64_alone(p, log_theta[p], nlml + p, grad ? grad[p] : nullptr, e) == GPRX_OK)
          h.theta.assign(log_theta[p], log_theta[p] + P + 1);
        else
          nlml[p] = NAN;
        continue;
      }
      if (B->hinfo.p[p]) {
        set_error(e, (B->hinfo.p[p] & 2) ? GPRX_NON_FINITE_KERNEL_VALUE : GPRX_CHOLESKY_FAILED,
                  "factor failed", GPRX_STAGE_FIT, 0.0, (size_t)h.n);
        nlml[p] = NAN;
        continue;
      }
      const double* sc = B->hscal.p + 8 * p;
      nlml[p] = 0.5 * sc[0] + sc[1] + 0.5 * h.n * kLog2Pi;
      if (grad && grad[p]) {
        acc_to_grad(h.hp, log_theta[p], B->hacc.p + (long)p * kMaxAcc, grad[p]);
        for (int q = 0; q < P; ++q) grad[p][q] *= -0.5;
        grad[p][P] = -0.5 * std::exp(log_theta[p][P]) * sc[2];
      }
      h.theta.assign(log_theta[p], log_theta[p] + P + 1);
    }
    return GPRX_OK;
  });
}

int gprx_hip_batch_set_schedule(gprx_hip_model* model, int32_t policy, int32_t small_tile_below,
                                gprx_error* err) {
  return guarded(err, [&]() -> int {
    if (!model || model->kind != gprx_hip_model::kBatch || policy < -1 || policy > 2) {
      set_error(err, GPRX_UNSUPPORTED, "bad handle or schedule policy");
      return GPRX_UNSUPPORTED;
    }
    auto* B = static_cast<BatchModel*>(model);
    // a captured evaluation holds the old slots: capture again (once no
    // launch of the old graphs is in flight)
    GPRX_HIP_CHECK(hipStreamSynchronize(B->stream));
    if (B->graph_exec) {
      GPRX_HIP_CHECK(hipGraphExecDestroy(B->graph_exec));
      B->graph_exec = nullptr;
    }
    if (B->graph) {
      GPRX_HIP_CHECK(hipGraphDestroy(B->graph));
      B->graph = nullptr;
    }
    if (B->graph_b_exec) {
      GPRX_HIP_CHECK(hipGraphExecDestroy(B->graph_b_exec));
      B->graph_b_exec = nullptr;
    }
    if (B->graph_b) {
      GPRX_HIP_CHECK(hipGraphDestroy(B->graph_b));
      B->graph_b = nullptr;
    }
    if (policy < 0) {
      B->policy = -1;
      B->forced = false;
      return GPRX_OK;
    }
    B->set_schedule(policy, small_tile_below < 0 ? 0 : small_tile_below);
    B->forced = true;
    B->tuned_ms = 0.0;
    return GPRX_OK;
  });
}

int gprx_hip_batch_schedule(const gprx_hip_model* model, int32_t* policy, int32_t* small_tile_below,
                            double* tuned_ms) {
  if (!model || model->kind != gprx_hip_model::kBatch) return GPRX_UNSUPPORTED;
  const auto* B = static_cast<const BatchModel*>(model);
  if (policy) *policy = B->policy;
  if (small_tile_below) *small_tile_below = B->small_below;
  if (tuned_ms) *tuned_ms = B->tuned_ms;
  return GPRX_OK;
}

int gprx_hip_batch_last_fallback(const gprx_hip_model* model, int32_t* reasons, size_t k) {
  if (!model || model->kind != gprx_hip_model::kBatch) return GPRX_UNSUPPORTED;
  const auto* B = static_cast<const BatchModel*>(model);
  if (!reasons || k != (size_t)B->K) return GPRX_LENGTH_MISMATCH;
  for (int p = 0; p < B->K; ++p) reasons[p] = B->fb_reason[p];
  return GPRX_OK;
}

}  // extern "C"
