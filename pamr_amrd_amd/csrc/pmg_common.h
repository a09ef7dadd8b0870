// pmg_common.h -- shared device helpers for the gfx950 multigrid kernels.
//
// Data layout: every grid function is a dense array in Fortran order
// (x fastest), exactly as PAMR stores grid functions (src/gh.c:884-900).
// Device code is compiled with -ffp-contract=off so that every stencil
// evaluates the same IEEE operations, in the same order, as the reference's
// Fortran (no FMA contraction): the FP64 path is bit-exact with it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#define PMG_HOST_CHECK(expr)                                                   \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "pamr_mg: %s failed: %s (%s:%d)\n", #expr,          \
                   hipGetErrorString(e_), __FILE__, __LINE__);                 \
      return -1;                                                               \
    }                                                                          \
  } while (0)

namespace pmg {

// Two HIP events for host-side timing, destroyed on every exit path
// (early returns of PMG_HOST_CHECK included).
struct EventPair {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t create() {
    hipError_t r = hipEventCreate(&e0);
    return r != hipSuccess ? r : hipEventCreate(&e1);
  }
  EventPair() = default;
  EventPair(const EventPair &) = delete;
  EventPair &operator=(const EventPair &) = delete;
  ~EventPair() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

constexpr int kWave = 64;        // CDNA wavefront width
constexpr int kXcd = 8;          // MI355X: 8 XCDs, each with its own L2

// The dispatcher hands workgroup b to XCD b % 8.  Map workgroups to tiles
// so that each XCD owns one contiguous range of tile indices (a bijection
// for any count): tiles that are neighbours in the linear tile order then
// share an L2, and halo re-reads hit it instead of going to HBM.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q = nb / kXcd, r = nb % kXcd;
  const int x = b % kXcd, local = b / kXcd;
  return x * q + (x < r ? x : r) + local;
}
constexpr int kMaxLev = 16;

__host__ __device__ inline size_t ix3(int i, int j, int k, int nx, int ny) {
  return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
}

// Scalars a stencil needs.  inv_* are used instead of a division only when
// the spacing squared is an exact power of two, where x*inv == x/d2 holds
// bit for bit (the unit-cube vertex grids, dx = 2^-m).
template <typename T>
struct Spacing {
  T dx2, dy2, dz2;
  T idx2, idy2, idz2;
  T jac;     // -2/dx2 - 2/dy2 - 2/dz2 (Poisson Jacobian, host-computed)
  T ijac;    // 1/jac (the FP32 inner cycles multiply instead of dividing)
  int pow2;  // 1 when inv_* are exact
};

// Wave64 + LDS block reduction of (sum, count) pairs; result valid in
// thread 0.  Fixed tree order: deterministic for a fixed launch shape.
template <int NT>
__device__ inline void block_reduce2(double &s, double &c) {
  __shared__ double red_s[NT / kWave], red_c[NT / kWave];
  for (int off = kWave / 2; off > 0; off >>= 1) {
    s += __shfl_down(s, off, kWave);
    c += __shfl_down(c, off, kWave);
  }
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (lane == 0) { red_s[wid] = s; red_c[wid] = c; }
  __syncthreads();
  if (wid == 0) {
    s = lane < NT / kWave ? red_s[lane] : 0.0;
    c = lane < NT / kWave ? red_c[lane] : 0.0;
    for (int off = kWave / 2; off > 0; off >>= 1) {
      s += __shfl_down(s, off, kWave);
      c += __shfl_down(c, off, kWave);
    }
  }
  __syncthreads();
}

// Norms of the reference hooks are RMS values sqrt(sum res^2 / count)
// (w/examples/nps/num_w.f:88,154).  A norm-producing kernel stores one
// (sum, count) partial per block in its slot of the workspace -- plain
// stores, no atomics and no cross-block fences -- and k_norm_fold, launched
// afterwards on the same stream, folds the partials of one or more slots in
// block order.  The result is deterministic for a fixed launch shape.
struct NormWs {
  double *partials;  // [nslots][cap][2]
  int *slot_nb;      // blocks that wrote each slot in its last launch
  int cap;           // max blocks per launch
  int nslots;
};

template <int NT>
__device__ inline void norm_partial(double s, double c, NormWs ws, int slot, int bid, int nb) {
  block_reduce2<NT>(s, c);
  if (threadIdx.x == 0) {
    double *p = ws.partials + ((size_t)slot * ws.cap + bid) * 2;
    p[0] = s;
    p[1] = c;
    if (bid == 0) ws.slot_nb[slot] = nb;
  }
}

// block b folds slots slot0 + b*fold ... slot0 + b*fold + fold-1 into out[b]
constexpr int kFoldThreads = 1024;
__global__ __launch_bounds__(kFoldThreads) void k_norm_fold(NormWs ws, int slot0, int fold, double *out);

}  // namespace pmg
