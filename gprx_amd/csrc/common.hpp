// Shared definitions of the gfx950 backend: the device form of a kernel
// program, status handling, and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gprx_hip.h"

namespace gprx {

// ---------------------------------------------------------------------------
// Kernel program in sum-of-products form.
//
// The C ABI takes gprx's KernelSpec tree (src/kernel/spec.rs:106-132) as a
// postfix list.  The host expands it into a sum of products of leaves,
//     K(a, b) = sum_t prod_{l in term t} v_l(a, b),
// and lays the factors out term by term as "slots": slot s holds one leaf
// occurrence, the slots of a term are contiguous, and bitmasks mark where
// terms open and close.  A leaf shared by several terms ((k1 + k2) * k3)
// occupies one slot per term.  Per matrix entry the kernels then need one
// forward sweep (running product of the open term: K) and, for gradients,
// one backward sweep (each slot's full term product, d v_s / d log theta_s =
// v_s f_s, so d K / d log theta_s = term product x f_s), every index known
// at compile time (registers only, no scratch).  Gradients of a leaf add up
// over its slots on the host.
// ---------------------------------------------------------------------------
constexpr int kMaxLeaves = 8;
constexpr int kMaxTerms = 8;
constexpr int kMaxSlots = 8;
constexpr int kMaxDim = 32;
constexpr int kTile = 128;  // tile edge of every dense kernel

enum LeafKind : int32_t { kLeafConst = 0, kLeafRbf = 1, kLeafArd = 2, kLeafPeriodic = 3 };

// Contraction accumulators: two per slot plus one per ARD dimension (the
// single ARD leaf of a program, summed over its slots).
constexpr int kAccPerLeaf = 2;
constexpr int kMaxAcc = kMaxSlots * kAccPerLeaf + kMaxDim;

// Index of the calling wave in its workgroup, read through the scalar unit:
// the compiler cannot prove threadIdx.x >> 6 uniform, and branches or
// register indices that depend on a value it takes as divergent become
// exec-mask control flow and per-lane selects.
__device__ __forceinline__ int wave_index() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// Sum over the 64 lanes of a wavefront by butterfly shuffles; every lane
// ends with the same value, added in the same order.
template <typename Op>
__device__ __forceinline__ double wave_reduce(double v, Op op) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
  return v;
}
struct SumOp {
  __device__ double operator()(double a, double b) const { return a + b; }
};
struct MaxOp {
  __device__ double operator()(double a, double b) const { return fmax(a, b); }
};
__device__ __forceinline__ double wave_sum(double v) { return wave_reduce(v, SumOp{}); }

// Reduction over a workgroup of NT threads (all must call it): each
// wavefront reduces by shuffles, the NT / 64 results meet in red (NT / 64
// doubles, one array per concurrent reduction) and wave 0 reduces them by
// shuffles again.  The result is in every lane of wave 0; the order is
// fixed, so a rerun gives the same bits.
template <int NT, typename Op>
__device__ __forceinline__ double block_reduce(double v, double* red, Op op, double init) {
  constexpr int NW = NT / 64;
  static_assert(NT % 64 == 0 && NW <= 64, "whole wavefronts, at most 64");
  v = wave_reduce(v, op);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = init;
  if (wave == 0) t = wave_reduce(lane < NW ? red[lane] : init, op);
  return t;
}
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  return block_reduce<NT>(v, red, SumOp{}, 0.0);
}
template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
  return block_reduce<NT>(v, red, MaxOp{}, 0.0);
}

struct DevProgram {
  int32_t ns;       // slots
  int32_t d;        // true input dimension (<= padded dimension)
  uint32_t start;   // bit s: slot s opens a term
  uint32_t end;     // bit s: slot s closes a term
  uint32_t ard;     // bit s: slot s holds the ARD leaf
  int32_t kind[kMaxSlots];
  double a[kMaxSlots];     // const: c; rbf: 1/l^2; periodic: pi/p
  double b[kMaxSlots];     // periodic: 1/l^2
  double ard_w[kMaxDim];   // 1/l_d^2 of the ARD leaf, 0 beyond d
  double noise;            // sigma^2 + jitter on the training diagonal
  // the same constants in f32, for entries evaluated in f32
  float af[kMaxSlots], bf[kMaxSlots], ard_wf[kMaxDim];
};

// Host-side description of a program: leaves in depth-first order, terms,
// and the slot layout.
struct HostProgram {
  int nl = 0, nt = 0, d = 0, ard_leaf = -1;
  int kind[kMaxLeaves];
  int param_off[kMaxLeaves];  // first theta index of each leaf
  std::vector<uint32_t> terms;
  int ns = 0;
  int slot_leaf[kMaxSlots];
  uint32_t start = 0, end = 0;
  int n_params = 0;  // kernel parameters (theta without log sigma^2)
};

// Parses the postfix program (pairs of int32 op, arg) for input dimension d.
// Returns GPRX_OK or an error status with a message.
int parse_program(const gprx_kernel_program* prog, int d, HostProgram* out, std::string* why);

// Fills a DevProgram from log-theta (kernel part only) and the diagonal term.
void make_dev_program(const HostProgram& hp, const double* log_theta, double diag_noise,
                      DevProgram* out);

// Number of contraction accumulators used by a program (2 per slot + d).
inline int n_acc(const HostProgram& hp) { return hp.ns * kAccPerLeaf + (hp.ard_leaf >= 0 ? hp.d : 0); }

// Maps contraction accumulators (sums of G_l * f) into d<sum W dK>/d(log theta),
// kernel parameters only.  acc has n_acc(hp) entries.
void acc_to_grad(const HostProgram& hp, const double* log_theta, const double* acc, double* grad);

// Padded input dimension used by the fill kernels (1, 2, 4, 8, 16 or 32).
inline int padded_dim(int d) {
  int p = 1;
  while (p < d) p *= 2;
  return p;
}

inline size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

// ---------------------------------------------------------------------------
// Status plumbing.
// ---------------------------------------------------------------------------
void set_error(gprx_error* err, int status, const std::string& msg, int stage = 0,
               double jitter = 0.0, size_t matrix_size = 0);

#define GPRX_HIP_CHECK(expr)                                                             \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) {                                                              \
      throw ::gprx::DeviceError(std::string(#expr) + ": " + hipGetErrorString(_e));      \
    }                                                                                    \
  } while (0)

struct DeviceError {
  std::string msg;
  explicit DeviceError(std::string m) : msg(std::move(m)) {}
};

// A device allocation that frees itself.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    GPRX_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Pinned host staging buffer.
template <typename T>
struct HostBuf {
  T* p = nullptr;
  size_t n = 0;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    GPRX_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocDefault));
    n = count;
  }
};

}  // namespace gprx
