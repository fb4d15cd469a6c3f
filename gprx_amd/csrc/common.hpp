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
// postfix list.  The host expands it into a sum of products of leaves, which
// the fill and contraction kernels evaluate per matrix entry with every index
// known at compile time (registers only, no scratch):
//     K(a, b) = sum_t prod_{l in term t} v_l(a, b)
// A leaf may appear in several terms after expansion ((k1 + k2) * k3); its
// parameters still live once in theta, so gradients add up across terms.
// ---------------------------------------------------------------------------
constexpr int kMaxLeaves = 8;
constexpr int kMaxTerms = 8;
constexpr int kMaxDim = 32;
constexpr int kTile = 128;  // tile edge of every dense kernel

enum LeafKind : int32_t { kLeafConst = 0, kLeafRbf = 1, kLeafArd = 2, kLeafPeriodic = 3 };

// Per-leaf slots of the contraction accumulators: two per leaf plus one per
// ARD dimension (the single ARD leaf of a program).
constexpr int kAccPerLeaf = 2;
constexpr int kMaxAcc = kMaxLeaves * kAccPerLeaf + kMaxDim;

struct DevProgram {
  int32_t nl;        // leaves
  int32_t nt;        // terms
  int32_t d;         // true input dimension (<= padded dimension)
  int32_t ard_leaf;  // the ARD leaf, or -1
  int32_t kind[kMaxLeaves];
  uint32_t mask[kMaxTerms];  // bit l set: leaf l is a factor of term t
  double a[kMaxLeaves];      // const: c; rbf: 1/l^2; periodic: pi/p
  double b[kMaxLeaves];      // periodic: 1/l^2
  double ard_w[kMaxDim];     // 1/l_d^2 of the ARD leaf, 0 beyond d
  double noise;              // sigma^2 + jitter on the training diagonal
};

// Host-side description of a program: leaves in depth-first order, terms.
struct HostProgram {
  int nl = 0, nt = 0, d = 0, ard_leaf = -1;
  int kind[kMaxLeaves];
  int param_off[kMaxLeaves];  // first theta index of each leaf
  std::vector<uint32_t> terms;
  int n_params = 0;  // kernel parameters (theta without log sigma^2)
};

// Parses the postfix program (pairs of int32 op, arg) for input dimension d.
// Returns GPRX_OK or an error status with a message.
int parse_program(const gprx_kernel_program* prog, int d, HostProgram* out, std::string* why);

// Fills a DevProgram from log-theta (kernel part only) and the diagonal term.
void make_dev_program(const HostProgram& hp, const double* log_theta, double diag_noise,
                      DevProgram* out);

// Number of contraction accumulators used by a program (2 per leaf + d).
inline int n_acc(const HostProgram& hp) { return hp.nl * kAccPerLeaf + (hp.ard_leaf >= 0 ? hp.d : 0); }

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
