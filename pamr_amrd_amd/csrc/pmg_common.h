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

constexpr int kWave = 64;        // CDNA wavefront width
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

// Norm epilogue shared by every norm-producing kernel: each block stores its
// (sum res^2, count) partial; the last block to finish (ticket counter)
// folds the partials in block order and writes sqrt(sum/count) -- the RMS
// norm of the reference hooks (w/examples/nps/num_w.f:88,154).  The ticket
// is reset by that last block, so the same workspace serves the next launch
// and graph replays without a memset.
struct NormWs {
  double *partials;  // 2 * capacity doubles
  unsigned *ticket;
  int capacity;
};

template <int NT>
__device__ inline void norm_epilogue(double s, double c, NormWs ws, double *out, int slot,
                                     int nblocks_launch, int nfold, bool final_launch) {
  block_reduce2<NT>(s, c);
  __shared__ int is_last;
  if (threadIdx.x == 0) {
    ws.partials[2 * slot] = s;
    ws.partials[2 * slot + 1] = c;
    is_last = 0;
    if (final_launch) {
      __threadfence();
      unsigned t = atomicAdd(ws.ticket, 1u);
      is_last = (t == (unsigned)(nblocks_launch - 1));
    }
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  double ts = 0.0, tc = 0.0;
  for (int b = threadIdx.x; b < nfold; b += NT) {
    ts += ws.partials[2 * b];
    tc += ws.partials[2 * b + 1];
  }
  block_reduce2<NT>(ts, tc);
  if (threadIdx.x == 0) {
    if (out) out[0] = sqrt(ts / tc);
    atomicExch(ws.ticket, 0u);
  }
}

}  // namespace pmg
