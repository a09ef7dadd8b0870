// pmg_amr.hip -- batched multigrid over an AMR hierarchy (several grids per
// level) on one GPU.
//
// Host side (once per hierarchy):
//   * the MG hierarchy of PAMR_build_mgh (src/gh.c:1222-1365): coarsened
//     copies of the base grid below it, then the AMR levels, grids ordered by
//     insert_grid (src/gh.c:235-257) as build_sgh and compose_cgh apply it;
//   * the transfer lists of every communicate() the V-cycle performs
//     (src/pamr.c:533-593): owned segments (src/transfer.c:415-458),
//     complete segments (:465-504), their pairwise intersections
//     (:514-540), calc_delta / align_src_bbox / calc_gs_ibbox index
//     arithmetic and the interpolation buffer offsets of local_transfer
//     (:946-1163).  Grid segments keep physical bounding boxes and the
//     reference's fuzz comparisons (src/misc.c:14-37), so the lists equal the
//     reference's piece for piece;
//   * each list is then made conflict-free: walking it backwards, every
//     transfer keeps only the destination points no later transfer writes
//     (sequential execution lets the last write win), cut into disjoint index
//     boxes.  All boxes of one list run in one launch.
// Device side: every level is a pool (one allocation per grid function,
// grids at 64-element aligned offsets); a sweep, residual or L_op covers all
// grids of a level in one launch through a tile table; norms fold per grid
// and then take the max over grids (call_rr_app, amrd/util.c:73-103).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/pamr_amr_gpu.h"
#include "../../include/pamr_mg_gpu.h"
#include "pmg_kernels.h"

using namespace pmg;

namespace {

// ---------------------------------------------------------------------------
// device descriptors
// ---------------------------------------------------------------------------
struct GridGeo {
  size_t off;  // element offset in the level pool
  int nx, ny, nz;
};
struct BTile {
  int g, x0, y0, z0;
};
struct XJob {
  int sg, dg;          // source grid (source level), destination grid
  int lo[3], hi[3];    // destination index box, 0-based inclusive
  int sh[3];           // copy: src = dst + sh; restrict: fine = 2 dst + sh; interp: F = dst + sh
  int t0[3], tn[3];    // interp: the buffered temporary's first coarse point (0-based) and point count
};
struct XBlk {
  int job, first;
};

constexpr int kXT = 256;              // threads of a transfer block
constexpr int PTX = 32, PTY = 8, PTZ = 8;  // point tiles of residual / L_op
constexpr int kMaxFields = 5;

template <typename T>
struct XFields {
  const T *src[kMaxFields];
  T *dst[kMaxFields];
  T *alt[kMaxFields];
  int nf;
};

__device__ __forceinline__ bool job_point(const XJob &J, int idx, int &i, int &j, int &k) {
  const int bx = J.hi[0] - J.lo[0] + 1, by = J.hi[1] - J.lo[1] + 1, bz = J.hi[2] - J.lo[2] + 1;
  if (idx >= bx * by * bz) return false;
  i = J.lo[0] + idx % bx;
  const int t = idx / bx;
  j = J.lo[1] + t % by;
  k = J.lo[2] + t / by;
  return true;
}

// sync copies (local_transfer, delta = 0)
template <typename T>
__global__ __launch_bounds__(kXT) void k_xfer_copy(XFields<T> F, const GridGeo *gs, const GridGeo *gd,
                                                   const XJob *jobs, const XBlk *blks) {
  const XBlk b = blks[blockIdx.x];
  const XJob J = jobs[b.job];
  int i, j, k;
  if (!job_point(J, b.first + threadIdx.x, i, j, k)) return;
  const GridGeo S = gs[J.sg], D = gd[J.dg];
  const size_t qd = D.off + ix3(i, j, k, D.nx, D.ny);
  const size_t qs = S.off + ix3(i + J.sh[0], j + J.sh[1], k + J.sh[2], S.nx, S.ny);
  for (int f = 0; f < F.nf; f++) {
    const T v = F.src[f][qs];
    F.dst[f][qd] = v;
    if (F.alt[f]) F.alt[f][qd] = v;
  }
}

// inject (local_transfer, delta = 1): dmcopy3d / dmhwr3d / dmfwr3d at
// stride 2 (src/m_util_r8.f:20-50, 649-830); the point operators select the
// reduced stencils from the position inside the fine grid, as the Fortran does
template <typename T, int KIND>
__global__ __launch_bounds__(kXT) void k_xfer_restrict(XFields<T> F, const GridGeo *gs, const GridGeo *gd,
                                                       const XJob *jobs, const XBlk *blks) {
  const XBlk b = blks[blockIdx.x];
  const XJob J = jobs[b.job];
  int i, j, k;
  if (!job_point(J, b.first + threadIdx.x, i, j, k)) return;
  const GridGeo S = gs[J.sg], D = gd[J.dg];
  const size_t qd = D.off + ix3(i, j, k, D.nx, D.ny);
  const int i0 = 2 * i + J.sh[0] + 1, j0 = 2 * j + J.sh[1] + 1, k0 = 2 * k + J.sh[2] + 1;  // 1-based fine
  for (int f = 0; f < F.nf; f++) {
    const T *u1 = F.src[f] + S.off;
    T v;
    if constexpr (KIND == 1) v = u1[ix3(i0 - 1, j0 - 1, k0 - 1, S.nx, S.ny)];
    else if constexpr (KIND == 2) v = hw_point<T, false>(u1, nullptr, T(0), S.nx, S.ny, S.nz, i0, j0, k0);
    else v = fw_point<T>(u1, S.nx, S.ny, S.nz, i0, j0, k0);
    F.dst[f][qd] = v;
    if (F.alt[f]) F.alt[f][qd] = v;
  }
}

// interp (local_transfer, delta = -1): dminterp3d second order
// (src/m_util_r8.f:473-637, dvi2q1d :332-360) evaluated per fine point: x
// phase on coarse-anchored lines, then y, then z, with c0 = c1 = 1/2
template <typename T>
__global__ __launch_bounds__(kXT) void k_xfer_interp(XFields<T> F, const GridGeo *gs, const GridGeo *gd,
                                                     const XJob *jobs, const XBlk *blks) {
  const XBlk b = blks[blockIdx.x];
  const XJob J = jobs[b.job];
  int i, j, k;
  if (!job_point(J, b.first + threadIdx.x, i, j, k)) return;
  const GridGeo S = gs[J.sg], D = gd[J.dg];
  const size_t qd = D.off + ix3(i, j, k, D.nx, D.ny);
  const int Fx = i + J.sh[0], Fy = j + J.sh[1], Fz = k + J.sh[2];
  const int ai = Fx >> 1, ri = Fx & 1, aj = Fy >> 1, rj = Fy & 1, ak = Fz >> 1, rk = Fz & 1;
  const T c0 = T(1 - (1 * 1.0) / 2), c1 = T((1 * 1.0) / 2);
  for (int f = 0; f < F.nf; f++) {
    const T *uc = F.src[f] + S.off;
    auto R = [&](int a, int bb, int c) -> T { return uc[ix3(a, bb, c, S.nx, S.ny)]; };
    auto X = [&](int bb, int c) -> T { return ri == 0 ? R(ai, bb, c) : c0 * R(ai, bb, c) + c1 * R(ai + 1, bb, c); };
    auto Y = [&](int c) -> T { return rj == 0 ? X(aj, c) : c0 * X(aj, c) + c1 * X(aj + 1, c); };
    const T v = rk == 0 ? Y(ak) : c0 * Y(ak) + c1 * Y(ak + 1);
    F.dst[f][qd] = v;
    if (F.alt[f]) F.alt[f][qd] = v;
  }
}

// interp at fourth order: dminterp3d ord 4 (src/m_util_r8.f:473-637, dvi4q1d
// :65-137) on the transfer's buffered temporary (coarse points t0 .. t0+tn-1
// per axis, src/transfer.c:1007-1051), whose ends take the one-sided cells,
// evaluated per fine point in the reference's phase order: x on
// coarse-anchored lines, y for every x, then z; anchors are copied, never
// re-evaluated
template <typename T>
__global__ __launch_bounds__(kXT) void k_xfer_interp4(XFields<T> F, const GridGeo *gs, const GridGeo *gd,
                                                      const XJob *jobs, const XBlk *blks) {
  const XBlk b = blks[blockIdx.x];
  const XJob J = jobs[b.job];
  int i, j, k;
  if (!job_point(J, b.first + threadIdx.x, i, j, k)) return;
  const GridGeo S = gs[J.sg], D = gd[J.dg];
  const size_t qd = D.off + ix3(i, j, k, D.nx, D.ny);
  const int fx = i + J.sh[0] - 2 * J.t0[0], fy = j + J.sh[1] - 2 * J.t0[1], fz = k + J.sh[2] - 2 * J.t0[2];
  for (int f = 0; f < F.nf; f++) {
    const T *uc = F.src[f] + S.off;
    auto line = [&](int d, int fd, const auto &A) -> T {
      if ((fd & 1) == 0) return A(fd >> 1);
      return vi4_point(J.tn[d], 2, 1, 2 * J.tn[d] - 1, fd, A);
    };
    auto Y = [&](int c) -> T {
      auto X = [&](int bb) -> T {
        auto R = [&](int a) -> T { return uc[ix3(J.t0[0] + a, J.t0[1] + bb, J.t0[2] + c, S.nx, S.ny)]; };
        return line(0, fx, R);
      };
      return line(1, fy, X);
    };
    const T v = line(2, fz, Y);
    F.dst[f][qd] = v;
    if (F.alt[f]) F.alt[f][qd] = v;
  }
}

// compute_cgc fused (amrd/mg.c:323-355), for levels whose interp plan writes
// every fine point exactly once: the coarse f_res = (u - lop)/w0_i is formed
// where the interpolation reads it and the interpolated value is added to u
// (and its ping-pong twin) at once; the same operations in the same order as
// the three-step form, without storing either f_res
template <typename T, int ORD>
__global__ __launch_bounds__(kXT) void k_xfer_cgc(const T *__restrict__ cu, const T *__restrict__ clop,
                                                  T w0_i, T *fu, T *fua, const GridGeo *gs,
                                                  const GridGeo *gd, const XJob *jobs, const XBlk *blks) {
  const XBlk b = blks[blockIdx.x];
  const XJob J = jobs[b.job];
  int i, j, k;
  if (!job_point(J, b.first + threadIdx.x, i, j, k)) return;
  const GridGeo S = gs[J.sg], D = gd[J.dg];
  const size_t qd = D.off + ix3(i, j, k, D.nx, D.ny);
  auto C = [&](int a, int bb, int c) -> T {
    const size_t q = S.off + ix3(a, bb, c, S.nx, S.ny);
    return (cu[q] - clop[q]) / w0_i;
  };
  T v;
  if constexpr (ORD == 2) {
    const int Fx = i + J.sh[0], Fy = j + J.sh[1], Fz = k + J.sh[2];
    const int ai = Fx >> 1, ri = Fx & 1, aj = Fy >> 1, rj = Fy & 1, ak = Fz >> 1, rk = Fz & 1;
    const T c0 = T(1 - (1 * 1.0) / 2), c1 = T((1 * 1.0) / 2);
    auto X = [&](int bb, int c) -> T { return ri == 0 ? C(ai, bb, c) : c0 * C(ai, bb, c) + c1 * C(ai + 1, bb, c); };
    auto Y = [&](int c) -> T { return rj == 0 ? X(aj, c) : c0 * X(aj, c) + c1 * X(aj + 1, c); };
    v = rk == 0 ? Y(ak) : c0 * Y(ak) + c1 * Y(ak + 1);
  } else {
    const int fx = i + J.sh[0] - 2 * J.t0[0], fy = j + J.sh[1] - 2 * J.t0[1], fz = k + J.sh[2] - 2 * J.t0[2];
    auto line = [&](int d, int fd, const auto &A) -> double {
      if ((fd & 1) == 0) return A(fd >> 1);
      return vi4_point(J.tn[d], 2, 1, 2 * J.tn[d] - 1, fd, A);
    };
    auto Y = [&](int c) -> double {
      auto X = [&](int bb) -> double {
        auto R = [&](int a) -> double { return C(J.t0[0] + a, J.t0[1] + bb, J.t0[2] + c); };
        return line(0, fx, R);
      };
      return line(1, fy, X);
    };
    v = (T)line(2, fz, Y);
  }
  const T nu = fu[qd] + v;
  fu[qd] = nu;
  fua[qd] = nu;
}

// ---------------------------------------------------------------------------
// batched stencil kernels over all grids of a level
// ---------------------------------------------------------------------------
template <typename T>
struct SweepBatchP {
  const T *uin;
  T *uout;
  const T *rho, *rhs, *a, *c;
  const GridGeo *grids;
  const BTile *tiles;
  int zchunk;
  Spacing<T> sp;
  NormWs ws;
  int slot;
};

template <typename T, int OP, bool P2, bool RHS0, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_sweep_batch(SweepBatchP<T> B) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const BTile tl = B.tiles[t];
  const GridGeo G = B.grids[tl.g];
  SweepP<T> P;
  P.uin = B.uin + G.off;
  P.uout = B.uout + G.off;
  P.rho = B.rho + G.off;
  P.rhs = B.rhs + G.off;
  P.a = B.a ? B.a + G.off : nullptr;
  P.c = B.c ? B.c + G.off : nullptr;
  P.chr = nullptr;
  P.nx = G.nx; P.ny = G.ny; P.nz = G.nz;
  P.bx0 = P.by0 = P.bz0 = 1;
  P.bx1 = G.nx - 2; P.by1 = G.ny - 2; P.bz1 = G.nz - 2;
  P.zchunk = B.zchunk;
  P.sp = B.sp;
  P.ws = B.ws;
  P.slot = B.slot;
  P.nb = gridDim.x;
  sweep_tile<T, OP, false, P2, RHS0, TX, TY, NT>(P, tl.x0, tl.y0, tl.z0, min(tl.z0 + B.zchunk, G.nz - 1), t);
}

template <typename T>
struct PtsBatchP {
  const T *u, *rho, *rhs, *a, *c;
  T *res, *lop, *rhs_out;
  const uint8_t *child;
  const GridGeo *grids;
  const BTile *tiles;
  Spacing<T> sp;
  T w0_r;
  NormWs ws;
  int slot;
};

// residual hook (w/examples/nps/num_w.f:59-90) on the interior of every grid
template <typename T, int OP, bool P2, bool RHS0>
__global__ __launch_bounds__(PTX * PTY) void k_resid_batch(PtsBatchP<T> B) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const BTile tl = B.tiles[t];
  const GridGeo G = B.grids[tl.g];
  const int i = tl.x0 + (int)(threadIdx.x % PTX), j = tl.y0 + (int)(threadIdx.x / PTX);
  const size_t sy = G.nx, sz = (size_t)G.nx * G.ny;
  double s = 0.0, cnt = 0.0;
  if (i >= 1 && i <= G.nx - 2 && j >= 1 && j <= G.ny - 2) {
    const T *u = B.u + G.off, *rho = B.rho + G.off, *rhs = B.rhs + G.off;
    const T *a = B.a ? B.a + G.off : nullptr, *c = B.c ? B.c + G.off : nullptr;
    T *res = B.res + G.off;
#pragma unroll
    for (int kk = 0; kk < PTZ; kk++) {
      const int k = tl.z0 + kk;
      if (k < 1 || k > G.nz - 2) continue;
      const size_t p = ix3(i, j, k, G.nx, G.ny);
      const T L = lop_at<T, OP, P2>(u, rho, a, c, p, sy, sz, B.sp, nullptr);
      T r;
      if constexpr (RHS0) r = L - T(0); else r = L - rhs[p];
      res[p] = r;
      s += (double)r * (double)r;
      cnt += 1.0;
    }
  }
  norm_partial<PTX * PTY>(s, cnt, B.ws, B.slot, t, gridDim.x);
}

// L_op + compute_MG_rhs (amrd/mg.c:124-126, 271-304) with the child mask
template <typename T, int OP, bool P2>
__global__ __launch_bounds__(PTX * PTY) void k_lop_rhs_batch(PtsBatchP<T> B) {
  const int t = blockIdx.x;
  const BTile tl = B.tiles[t];
  const GridGeo G = B.grids[tl.g];
  const int i = tl.x0 + (int)(threadIdx.x % PTX), j = tl.y0 + (int)(threadIdx.x / PTX);
  if (i >= G.nx || j >= G.ny) return;
  const size_t sy = G.nx, sz = (size_t)G.nx * G.ny;
  const T *u = B.u + G.off, *rho = B.rho + G.off, *res = B.res + G.off;
  const T *a = B.a ? B.a + G.off : nullptr, *c = B.c ? B.c + G.off : nullptr;
  const uint8_t *child = B.child + G.off;
  T *lop = B.lop + G.off, *rhs = B.rhs_out + G.off;
  const bool ixy = i >= 1 && i <= G.nx - 2 && j >= 1 && j <= G.ny - 2;
#pragma unroll
  for (int kk = 0; kk < PTZ; kk++) {
    const int k = tl.z0 + kk;
    if (k >= G.nz) break;
    const size_t p = ix3(i, j, k, G.nx, G.ny);
    const bool on = child[p] != 0;
    T lopv = T(0);  // f_lop was cleared by clear_MG_vars
    if (on && ixy && k >= 1 && k <= G.nz - 2) lopv = lop_at<T, OP, P2>(u, rho, a, c, p, sy, sz, B.sp, nullptr);
    rhs[p] = on ? (-res[p]) * B.w0_r + lopv : T(0);
    lop[p] = u[p];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_add_alt(T *u, T *alt, const T *r, size_t n) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) {
    const T v = u[q] + r[q];
    u[q] = v;
    alt[q] = v;
  }
}

// mixed precision: the FP64 defect r = L[u] (the composite equation has rhs
// 0 on every AMR level's own points) of every grid of a level, written
// negated in FP32 as the source of the FP32 correction cycle (0 off the
// interior), with the RMS partials of the level's grids
template <int OP, bool P2>
__global__ __launch_bounds__(PTX * PTY) void k_defect32_batch(PtsBatchP<double> B, float *src32) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const BTile tl = B.tiles[t];
  const GridGeo G = B.grids[tl.g];
  const int i = tl.x0 + (int)(threadIdx.x % PTX), j = tl.y0 + (int)(threadIdx.x / PTX);
  const size_t sy = G.nx, sz = (size_t)G.nx * G.ny;
  double s = 0.0, cnt = 0.0;
  if (i < G.nx && j < G.ny) {
    const double *u = B.u + G.off, *rho = B.rho + G.off;
    const double *a = B.a ? B.a + G.off : nullptr, *c = B.c ? B.c + G.off : nullptr;
    float *out = src32 + G.off;
    const bool ixy = i >= 1 && i <= G.nx - 2 && j >= 1 && j <= G.ny - 2;
#pragma unroll
    for (int kk = 0; kk < PTZ; kk++) {
      const int k = tl.z0 + kk;
      if (k >= G.nz) break;
      const size_t p = ix3(i, j, k, G.nx, G.ny);
      float v = 0.0f;
      if (ixy && k >= 1 && k <= G.nz - 2) {
        const double r = lop_at<double, OP, P2>(u, rho, a, c, p, sy, sz, B.sp, nullptr) - 0.0;
        v = (float)(-r);
        s += r * r;
        cnt += 1.0;
      }
      out[p] = v;
    }
  }
  norm_partial<PTX * PTY>(s, cnt, B.ws, B.slot, t, gridDim.x);
}

// u += e in FP64 over a level's pool, both ping-pong buffers
__global__ __launch_bounds__(256) void k_add32_alt(double *u, double *ua, const float *e, size_t n) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) {
    const double v = u[q] + (double)e[q];
    u[q] = v;
    ua[q] = v;
  }
}

__global__ __launch_bounds__(256) void k_to_f32(float *dst, const double *src, size_t n) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) dst[q] = (float)src[q];
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill0(T *a, T *b, size_t n) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) {
    a[q] = T(0);
    b[q] = T(0);
  }
}

// per-grid RMS of the partials of one slot, then the max over the grids
// (call_rr_app: the first grid's value, then max)
struct FoldSpec {
  int slot;
  int ng;
  const int *gstart;  // ng+1 tile starts
};
// one block per norm slot: each wave sums the tile partials of every fourth
// grid (fixed lane order, then a fixed shuffle tree), the per-grid RMS values
// go to LDS, and one thread folds them with call_rr_app's max in grid order
constexpr int kFoldGrids = 512;  // grids folded through LDS; more are folded in chunks
__global__ __launch_bounds__(256) void k_fold_max(NormWs ws, const FoldSpec *specs, double *out) {
  const FoldSpec F = specs[blockIdx.x];
  const double *p = ws.partials + (size_t)F.slot * ws.cap * 2;
  __shared__ double rg[kFoldGrids];
  const int lane = (int)threadIdx.x & (kWave - 1), w = (int)threadIdx.x / kWave;
  double best = 0.0;
  for (int g0 = 0; g0 < F.ng; g0 += kFoldGrids) {
    const int ng = min(kFoldGrids, F.ng - g0);
    for (int q = w; q < ng; q += 256 / kWave) {
      const int g = g0 + q;
      double s = 0.0, c = 0.0;
      for (int b = F.gstart[g] + lane; b < F.gstart[g + 1]; b += kWave) {
        s += p[2 * b];
        c += p[2 * b + 1];
      }
      for (int off = kWave / 2; off > 0; off >>= 1) {
        s += __shfl_down(s, off, kWave);
        c += __shfl_down(c, off, kWave);
      }
      if (lane == 0) rg[q] = sqrt(s / c);
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int q = 0; q < ng; q++) {
        const double r = rg[q];
        best = g0 + q == 0 ? r : (r > best ? r : best);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = best;
}

// clear_MG_vars on one level: res, rhs and lop in one launch
// clear_ex (src/transfer.c:930-935): f = 0 where chr == ex
__global__ __launch_bounds__(256) void k_clear_ex(double *f, const double *chr, double ex, size_t n) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n && chr[q] == ex) f[q] = 0.0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_clear3(T *a, T *b, T *c, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  a[i] = T(0);
  b[i] = T(0);
  c[i] = T(0);
}

inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

// ---------------------------------------------------------------------------
// host bookkeeping (physical bounding boxes, the reference's rules)
// ---------------------------------------------------------------------------
bool fz_eq(double a, double b, double f) { return std::fabs(b - a) < f; }
bool fz_lt(double a, double b, double f) { return a < -f + b; }
bool fz_gt(double a, double b, double f) { return a > f + b; }

struct HGrid {
  int n[3];
  double bbox[6];
  double hdx[3], gdx[3];
  bool coarsest = false;
  size_t off = 0;
  double dx(int d) const { return n[d] > 1 ? (bbox[2 * d + 1] - bbox[2 * d]) / (n[d] - 1) : 1e-20; }
  size_t size() const { return (size_t)n[0] * n[1] * n[2]; }
};

struct Seg {
  int g;
  double bb[6];
};

// calc_gs_ibbox (src/transfer.c:73-95): 1-based index box of a segment
void seg_ibbox(const Seg &s, const HGrid &g, int *ib, double *dx) {
  for (int i = 0; i < 3; i++) {
    if (g.n[i] > 1) {
      dx[i] = (g.bbox[2 * i + 1] - g.bbox[2 * i]) / (g.n[i] - 1);
      ib[2 * i] = (int)((s.bb[2 * i] - g.bbox[2 * i]) / dx[i] + 1.01);
      ib[2 * i + 1] = (int)((s.bb[2 * i + 1] - g.bbox[2 * i]) / dx[i] + 1.99);
    } else {
      ib[2 * i] = ib[2 * i + 1] = 1;
      dx[i] = 1e-20;
    }
  }
}

struct Hier;

// ---------------------------------------------------------------------------
// level state
// ---------------------------------------------------------------------------
struct XPlan {
  std::vector<XJob> raw;   // the transfer list in the reference's order (excision path)
  std::vector<XJob> jobs;  // resolved into disjoint boxes (batched path)
  XJob *d_jobs = nullptr;
  XBlk *d_blks = nullptr;
  int nblocks = 0;
};

struct ALevel {
  std::vector<HGrid> g;
  bool in_amrh = false;
  size_t pool = 0;
  double *u = nullptr, *ua = nullptr, *rho = nullptr, *rhs = nullptr, *res = nullptr, *lop = nullptr;
  double *a = nullptr, *c = nullptr;
  std::vector<double *> extra;  // fields added by pmg_amr_add_field
  double *cm[2] = {nullptr, nullptr};  // set_cmask_bdy / set_cmask_child as doubles (built on first use)
  uint8_t *child = nullptr;
  GridGeo *d_geo = nullptr;
  Spacing<double> sp;
  Spacing<float> sp32;
  // FP32 shadows for the inner cycles of mixed precision (allocated on first
  // use), indexed like the fields: u, rho, rhs, res, lop, a, c, then ua
  float *f32[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int variant = 2, zchunk = 16;
  BTile *d_stiles = nullptr;
  int nstiles = 0;
  int *d_sstart = nullptr;
  BTile *d_ptiles = nullptr;
  int nptiles = 0;
  int *d_pstart = nullptr;
  XPlan sync0, sync1, inject, interp;  // inject: from L+1 into L; interp: from L-1 into L
  XPlan sync1c;  // sync (AMR_bdy_width 1) among the coarsest grids only (comm = coarsest)
  bool interp_once = false;  // the interp plan writes every point of every grid exactly once
  bool shared = false;       // grid functions belong to the parent hierarchy, its level plev
  int plev = 0;
};

}  // namespace

struct pmg_amr {
  int nlev = 0, nbase = 0, op = 0;
  int nextra = 0;
  double cbbox[6];
  int periodic[3] = {0, 0, 0};
  std::vector<ALevel> lev;
  int pre = 2, pst = 2, max_citer = 5000, rop_u = 3, rop_res = 3, rop_src = 3, iord = 2;
  double crtol = 1e-3, w0_r = 1.0, w0_i = 1.0;
  hipStream_t stream = nullptr;
  NormWs ws{nullptr, nullptr, 0, 0};
  double *d_norms = nullptr;  // [0, kMaxLev): relax per level; [kMaxLev, 2 kMaxLev): residual; 2 kMaxLev: final
  double *d_cnorm = nullptr;  // coarse solve pre / post
  double *d_snap = nullptr;   // saved state of level 1 for the coarsest loop's cycle watch
  int *d_citer = nullptr;
  FoldSpec *d_fold = nullptr;
  int nfold = 0;
  double *h_pinned = nullptr;
  int use_graph = 1;
  hipGraphExec_t graph = nullptr;
  hipGraphExec_t graph_res = nullptr;
  hipGraphExec_t graph_mixed = nullptr;  // one mixed-precision iteration
  bool have_fp32 = false;
  double mixed_switch = 1e-2;  // mode 1: FP64 V-cycles while ||r|| > mixed_switch ||r0||
  double *xtmp = nullptr;  // interpolation temporary of the excision path
  size_t xtmp_cap = 0;
  // the boxes given for each AMR level (index l-1) in the caller's order
  // (save_sgh_bboxes, src/gh.c:273-298); level 1 holds the base box
  std::vector<std::vector<double>> sgh;
  double dx0[3] = {0, 0, 0};
  int mcw = 5;
  const pmg_amr *parent = nullptr;  // an MG hierarchy over AMR levels of parent
  bool own_stream = true;
};

namespace {

enum { AS_RELAX = 0, AS_RESID = kMaxLev, AS_FINAL = 2 * kMaxLev, kAmrSlots = 2 * kMaxLev + 2 };

// alloc_gsl with ghost_width 0 (src/transfer.c:132-188)
bool seg_alloc(const pmg_amr *A, const HGrid &g, int gi, int interior, int AMR_reduce, Seg &out) {
  out.g = gi;
  for (int d = 0; d < 3; d++) {
    const double dx = g.dx(d);
    out.bb[2 * d] = g.bbox[2 * d];
    out.bb[2 * d + 1] = g.bbox[2 * d + 1];
    if (AMR_reduce > 0) {
      if (fz_gt(out.bb[2 * d], A->cbbox[2 * d], dx / 2) || A->periodic[d]) out.bb[2 * d] += AMR_reduce * dx;
      if (fz_lt(out.bb[2 * d + 1], A->cbbox[2 * d + 1], dx / 2) || A->periodic[d]) out.bb[2 * d + 1] += -AMR_reduce * dx;
    }
    if (interior || AMR_reduce) {
      int ib[6];
      double dxt[3];
      seg_ibbox(out, g, ib, dxt);
      if ((ib[2 * d + 1] - ib[2 * d] + 1) < 1) return false;
    }
  }
  return true;
}

// gs_subtract (src/transfer.c:218-283); pieces appended in the order of the
// reference's result list (each new piece is prepended there)
void seg_subtract(const std::vector<HGrid> &grids, const Seg &A, const Seg &B, std::vector<Seg> &out) {
  double cut[6], dx[3];
  const HGrid &g = grids[A.g];
  for (int i = 0; i < 3; i++) {
    cut[2 * i] = A.bb[2 * i];
    cut[2 * i + 1] = A.bb[2 * i + 1];
    dx[i] = g.dx(i);
    if (fz_gt(B.bb[2 * i], A.bb[2 * i + 1], dx[i] / 2) || fz_gt(A.bb[2 * i], B.bb[2 * i + 1], dx[i] / 2)) {
      out.push_back(A);
      return;
    }
  }
  std::vector<Seg> made;
  for (int i = 0; i < 3; i++) {
    cut[2 * i] = std::max(A.bb[2 * i], B.bb[2 * i]);
    if (fz_gt(cut[2 * i], A.bb[2 * i], dx[i] / 2)) {
      Seg C{A.g, {}};
      for (int j = 0; j < 3; j++) {
        if (i == j) {
          C.bb[2 * i] = A.bb[2 * i];
          C.bb[2 * i + 1] = std::max(C.bb[2 * i], cut[2 * i] - dx[i]);
        } else {
          C.bb[2 * j] = cut[2 * j];
          C.bb[2 * j + 1] = cut[2 * j + 1];
        }
      }
      made.push_back(C);
    }
    cut[2 * i + 1] = std::min(A.bb[2 * i + 1], B.bb[2 * i + 1]);
    if (fz_lt(cut[2 * i + 1], A.bb[2 * i + 1], dx[i] / 2)) {
      Seg C{A.g, {}};
      for (int j = 0; j < 3; j++) {
        if (i == j) {
          C.bb[2 * i + 1] = A.bb[2 * i + 1];
          C.bb[2 * i] = std::min(C.bb[2 * i + 1], cut[2 * i + 1] + dx[i]);
        } else {
          C.bb[2 * j] = cut[2 * j];
          C.bb[2 * j + 1] = cut[2 * j + 1];
        }
      }
      made.push_back(C);
    }
  }
  for (auto it = made.rbegin(); it != made.rend(); ++it) out.push_back(*it);
}

// build_owned_gsl (src/transfer.c:415-458) for one rank, with extend_gsl
std::vector<Seg> owned_segs(const pmg_amr *A, int L, int interior, int extend, bool coarsest_only = false) {
  const std::vector<HGrid> &grids = A->lev[L - 1].g;
  const int cint = interior ? 1 : 0, cred = interior > 1 ? interior - 1 : 0;
  std::vector<Seg> owned, prior;
  for (int gi = 0; gi < (int)grids.size(); gi++) {
    Seg one;
    if (coarsest_only && !grids[gi].coarsest) continue;  // comm == 0 (src/transfer.c:438)
    if (!seg_alloc(A, grids[gi], gi, cint, cred, one)) continue;
    std::vector<Seg> C{one};
    for (const Seg &B : prior) {  // gsl_subtract (src/transfer.c:285-314)
      std::vector<Seg> C0;
      for (const Seg &c : C) seg_subtract(grids, c, B, C0);
      C.swap(C0);
    }
    owned.insert(owned.end(), C.begin(), C.end());
    prior.push_back(one);
  }
  if (extend)
    for (Seg &s : owned) {
      const HGrid &g = grids[s.g];
      for (int d = 0; d < 3; d++)
        if (g.n[d] > 1) {
          const double dx = g.dx(d);
          s.bb[2 * d] = std::max(s.bb[2 * d] - dx, g.bbox[2 * d]);
          s.bb[2 * d + 1] = std::min(s.bb[2 * d + 1] + dx, g.bbox[2 * d + 1]);
        }
    }
  return owned;
}

// communicate() transfer list -> XJobs (before conflict resolution)
int make_jobs(const pmg_amr *A, int l1, int l2, int interior, int extend, std::vector<XJob> &jobs,
              bool coarsest_only = false) {
  const std::vector<HGrid> &G1 = A->lev[l1 - 1].g, &G2 = A->lev[l2 - 1].g;
  std::vector<Seg> src = owned_segs(A, l1, interior, extend, coarsest_only), dst;
  for (int gi = 0; gi < (int)G2.size(); gi++) {  // build_complete_gsl, AMR_bdy_only = 0
    Seg one;
    if (coarsest_only && !G2[gi].coarsest) continue;  // comm == 0 (src/transfer.c:482)
    if (seg_alloc(A, G2[gi], gi, 0, 0, one)) dst.push_back(one);
  }
  for (const Seg &s : src)
    for (const Seg &d : dst) {  // build_gstl + gs_and
      if (l1 == l2 && s.g == d.g) continue;
      const HGrid &ga = G1[s.g], &gb = G2[d.g];
      double nb[6];
      bool ok = true;
      for (int i = 0; i < 3; i++) {
        const double dxm = std::max(ga.dx(i), gb.dx(i));
        nb[2 * i] = std::max(s.bb[2 * i], d.bb[2 * i]);
        nb[2 * i + 1] = std::min(s.bb[2 * i + 1], d.bb[2 * i + 1]);
        if (fz_lt(nb[2 * i + 1], nb[2 * i], dxm / 2)) ok = false;
      }
      if (!ok) continue;
      Seg ss{s.g, {}}, dd{d.g, {}};
      std::memcpy(ss.bb, nb, sizeof nb);
      std::memcpy(dd.bb, nb, sizeof nb);
      // calc_delta (src/transfer.c:545-575)
      double dxsm = 0, dxdm = 0;
      for (int i = 0; i < 3; i++)
        if (ga.n[i] > 1 && gb.n[i] > 1) {
          dxsm = std::max(dxsm, ga.dx(i));
          dxdm = std::max(dxdm, gb.dx(i));
        }
      int delta = 0, irho = 0;
      if (dxsm == 0 || dxdm == 0 || fz_eq(dxsm, dxdm, dxsm / 100)) delta = 0;
      else if (fz_gt(dxsm, dxdm, dxsm / 100)) { delta = -1; irho = (int)(dxsm / dxdm + 0.5); }
      else { delta = 1; irho = (int)(dxdm / dxsm + 0.5); }
      if (delta != 0 && irho != 2) return -6;
      if (delta == 1) {  // align_src_bbox (src/transfer.c:101-121)
        for (int i = 0; i < 3; i++) {
          const double dxs = ga.dx(i), dxd = gb.dx(i);
          ss.bb[2 * i] = ((int)((ss.bb[2 * i] - gb.bbox[2 * i] + dxs / 2) / dxd)) * dxd + gb.bbox[2 * i];
          if (fz_lt(ss.bb[2 * i], ga.bbox[2 * i], dxs / 2)) ss.bb[2 * i] += dxd;
          dd.bb[2 * i] = ss.bb[2 * i];
          ss.bb[2 * i + 1] = ((int)((ss.bb[2 * i + 1] - gb.bbox[2 * i] + dxd - dxs / 2) / dxd)) * dxd + gb.bbox[2 * i];
          if (fz_gt(ss.bb[2 * i + 1], ga.bbox[2 * i + 1], dxs / 2)) ss.bb[2 * i + 1] -= dxd;
          dd.bb[2 * i + 1] = ss.bb[2 * i + 1];
        }
      }
      int ib[6], ibd[6];
      double dx[3], dxd[3];
      seg_ibbox(ss, ga, ib, dx);
      seg_ibbox(dd, gb, ibd, dxd);
      XJob J;
      J.sg = s.g;
      J.dg = d.g;
      bool empty = false;
      for (int i = 0; i < 3; i++) {
        const int cnt = delta == 0 ? ib[2 * i + 1] - ib[2 * i] + 1 : ibd[2 * i + 1] - ibd[2 * i] + 1;
        J.lo[i] = ibd[2 * i] - 1;
        J.hi[i] = J.lo[i] + cnt - 1;
        if (cnt < 1) empty = true;
        if (delta == 0) {
          J.sh[i] = ib[2 * i] - ibd[2 * i];
        } else if (delta == 1) {
          J.sh[i] = (ib[2 * i] - 1) - 2 * (ibd[2 * i] - 1);
        } else {
          // local_transfer's buffered interpolation (src/transfer.c:1007-1051):
          // destination point q reads the temporary at fine offset
          // b2 - 1 + (q - (ibd - 1)) from the coarse grid's first point
          const int b2 = (int)((ss.bb[2 * i] - ga.bbox[2 * i]) / dxd[i] + 1.5);
          J.sh[i] = b2 - ibd[2 * i];
          // the temporary: interp_buffer = 2 coarse points either side, clipped
          const int b2i = std::max(1, (((b2 - 1) / irho) - 2) * irho + 1), b1i = (b2i - 1) / irho + 1;
          J.t0[i] = b1i - 1;
          J.tn[i] = std::min(ga.n[i], (b2 + cnt - 1) / irho + 1 + 2 + 1) - b1i + 1;
        }
      }
      if (empty) continue;  // the Fortran loops run zero times
      // bounds of the reads (the device kernels do not check them)
      for (int i = 0; i < 3; i++) {
        const int nd = gb.n[i], ns = ga.n[i];
        if (J.lo[i] < 0 || J.hi[i] >= nd) return -7;
        if (delta == 0 && (J.lo[i] + J.sh[i] < 0 || J.hi[i] + J.sh[i] >= ns)) return -7;
        if (delta == 1 && (2 * J.lo[i] + J.sh[i] < 0 || 2 * J.hi[i] + J.sh[i] >= ns)) return -7;
        if (delta == -1) {
          const int f0 = J.lo[i] + J.sh[i], f1 = J.hi[i] + J.sh[i];
          if (f0 < 0 || (f1 >> 1) + ((f1 & 1) ? 1 : 0) >= ns) return -7;
          if (J.t0[i] < 0 || J.tn[i] < 1 || J.t0[i] + J.tn[i] > ns || f0 - 2 * J.t0[i] < 0 ||
              f1 - 2 * J.t0[i] > 2 * (J.tn[i] - 1))
            return -7;
        }
      }
      jobs.push_back(J);
    }
  return 0;
}

// A - B for integer boxes (inclusive), appended to out
void box_minus(const XJob &A, const int *blo, const int *bhi, std::vector<XJob> &out) {
  for (int d = 0; d < 3; d++)
    if (bhi[d] < A.lo[d] || blo[d] > A.hi[d]) {
      out.push_back(A);
      return;
    }
  XJob rest = A;
  for (int d = 0; d < 3; d++) {
    if (blo[d] > rest.lo[d]) {
      XJob p = rest;
      p.hi[d] = blo[d] - 1;
      out.push_back(p);
      rest.lo[d] = blo[d];
    }
    if (bhi[d] < rest.hi[d]) {
      XJob p = rest;
      p.lo[d] = bhi[d] + 1;
      out.push_back(p);
      rest.hi[d] = bhi[d];
    }
  }
}

// keep of every job the destination points no later job writes
std::vector<XJob> resolve(const std::vector<XJob> &jobs) {
  std::vector<XJob> out;
  std::vector<std::vector<XJob>> claimed;
  for (int q = (int)jobs.size() - 1; q >= 0; q--) {
    const XJob &J = jobs[q];
    if ((int)claimed.size() <= J.dg) claimed.resize(J.dg + 1);
    std::vector<XJob> pieces{J};
    for (const XJob &C : claimed[J.dg]) {
      std::vector<XJob> next;
      for (const XJob &p : pieces) box_minus(p, C.lo, C.hi, next);
      pieces.swap(next);
      if (pieces.empty()) break;
    }
    out.insert(out.end(), pieces.begin(), pieces.end());
    claimed[J.dg].push_back(J);
  }
  return out;
}

int upload_plan(pmg_amr *A, XPlan &P, const std::vector<XJob> &jobs) {
  P.raw = jobs;
  P.jobs = resolve(jobs);
  std::vector<XBlk> blks;
  for (int q = 0; q < (int)P.jobs.size(); q++) {
    const XJob &J = P.jobs[q];
    const long long n = (long long)(J.hi[0] - J.lo[0] + 1) * (J.hi[1] - J.lo[1] + 1) * (J.hi[2] - J.lo[2] + 1);
    for (long long f = 0; f < n; f += kXT) blks.push_back(XBlk{q, (int)f});
  }
  P.nblocks = (int)blks.size();
  if (P.jobs.empty()) return 0;
  PMG_HOST_CHECK(hipMalloc(&P.d_jobs, sizeof(XJob) * P.jobs.size()));
  PMG_HOST_CHECK(hipMemcpy(P.d_jobs, P.jobs.data(), sizeof(XJob) * P.jobs.size(), hipMemcpyHostToDevice));
  PMG_HOST_CHECK(hipMalloc(&P.d_blks, sizeof(XBlk) * blks.size()));
  PMG_HOST_CHECK(hipMemcpy(P.d_blks, blks.data(), sizeof(XBlk) * blks.size(), hipMemcpyHostToDevice));
  return 0;
}

// insert_grid (src/gh.c:235-257) position of a new bbox in a level list
void insert_sorted(std::vector<std::vector<double>> &list, const double *bb) {
  size_t p = 0;
  while (p < list.size() && list[p][0] < bb[0] && list[p][2] < bb[2] && list[p][4] < bb[4] && list[p][1] < bb[1] &&
         list[p][3] < bb[3] && list[p][5] < bb[5])
    p++;
  list.insert(list.begin() + p, std::vector<double>(bb, bb + 6));
}

void grid_geometry(HGrid &g, const double *bb, const double *ldx) {
  std::memcpy(g.bbox, bb, 6 * sizeof(double));
  for (int d = 0; d < 3; d++) {
    g.n[d] = (int)((bb[2 * d + 1] - bb[2 * d]) / ldx[d] + 1.5);
    const double dx = (bb[2 * d + 1] - bb[2 * d]) / (g.n[d] - 1);
    const double x0 = bb[2 * d], x1 = x0 + dx, x2 = x1 + dx;  // src/gh.c:905-906
    g.hdx[d] = x1 - x0;
    g.gdx[d] = x2 - x1;
  }
}

// build_sgh and compose_cgh each insert the grids with insert_grid
std::vector<HGrid> level_grids(int nb, const double *boxes, const double *ldx) {
  std::vector<std::vector<double>> l1, l2;
  for (int k = 0; k < nb; k++) insert_sorted(l1, boxes + 6 * k);
  for (auto &b : l1) insert_sorted(l2, b.data());
  std::vector<HGrid> out(nb);
  for (int k = 0; k < nb; k++) grid_geometry(out[k], l2[k].data(), ldx);
  return out;
}

// build_sgh + compose_cgh order of a box list, with perm[k] = the index in
// `boxes` of grid k
std::vector<HGrid> level_grids_perm(int nb, const double *boxes, const double *ldx, std::vector<int> &perm) {
  std::vector<std::vector<double>> l1, l2;
  std::vector<int> i1, i2;
  auto insert = [](std::vector<std::vector<double>> &list, std::vector<int> &idx, const double *bb, int id) {
    size_t p = 0;
    while (p < list.size() && list[p][0] < bb[0] && list[p][2] < bb[2] && list[p][4] < bb[4] && list[p][1] < bb[1] &&
           list[p][3] < bb[3] && list[p][5] < bb[5])
      p++;
    list.insert(list.begin() + p, std::vector<double>(bb, bb + 6));
    idx.insert(idx.begin() + p, id);
  };
  for (int k = 0; k < nb; k++) insert(l1, i1, boxes + 6 * k, k);
  for (int k = 0; k < nb; k++) insert(l2, i2, l1[k].data(), i1[k]);
  std::vector<HGrid> out(nb);
  for (int k = 0; k < nb; k++) grid_geometry(out[k], l2[k].data(), ldx);
  perm = i2;
  return out;
}

// calc_island_no (src/gh.c:1427-1467) of a box list: per box 0 when it (or a
// box it is joined with) cannot be coarsened by cf, negative when it becomes
// coarsest there.  Boxes are joined by the reference's own test (each
// dimension: not lo_i > hi_j and not hi_i > lo_j, with fuzz dx/2).
std::vector<int> island_numbers(const double *bb, int num, const double *dx, int cf, int mcw) {
  std::vector<int> isl(num);
  for (int i = 0; i < num; i++) {
    isl[i] = i + 1;
    for (int j = 0; j < 3; j++) {
      const int shape = (int)((bb[6 * i + 2 * j + 1] - bb[6 * i + 2 * j]) / dx[j] + 1.5);
      if ((((shape - 1) % cf) && ((shape - 1) / cf + 1) >= mcw) || (((shape - 1) / cf + 1) < mcw)) isl[i] = 0;
      else if (((shape - 1) / cf / 2 + 1) < mcw) isl[i] = -std::abs(isl[i]);
    }
  }
  for (int delta = 1; delta;) {
    delta = 0;
    for (int i = 0; i < num; i++)
      for (int j = i + 1; j < num; j++) {
        if (isl[i] == isl[j]) continue;
        bool overlap = true;
        for (int k = 0; k < 3; k++)
          if (fz_gt(bb[6 * i + 2 * k], bb[6 * j + 2 * k + 1], dx[k] / 2) ||
              fz_gt(bb[6 * i + 2 * k + 1], bb[6 * j + 2 * k], dx[k] / 2))
            overlap = false;
        if (overlap) {
          delta++;
          if (isl[i] == 0 || isl[j] == 0) isl[i] = isl[j] = 0;
          else isl[i] = isl[j] = std::min(isl[i], isl[j]);
        }
      }
  }
  return isl;
}

// calc_island_no (src/gh.c:1427-1467) of a single box
int island_single(const double *bb, const double *dx, int cf, int mcw) {
  int isl = 1;
  for (int j = 0; j < 3; j++) {
    const int shape = (int)((bb[2 * j + 1] - bb[2 * j]) / dx[j] + 1.5);
    if ((((shape - 1) % cf) && ((shape - 1) / cf + 1) >= mcw) || (((shape - 1) / cf + 1) < mcw)) isl = 0;
    else if (((shape - 1) / cf / 2 + 1) < mcw) isl = -std::abs(isl);
  }
  return isl;
}

bool is_pow2(double v) {
  int e;
  const double m = std::frexp(v, &e);
  return m == 0.5 && v > 1e-300 && v < 1e300;
}

Spacing<float> spacing_of32(const double *hdx) {
  Spacing<float> sp;
  sp.dx2 = (float)(hdx[0] * hdx[0]);
  sp.dy2 = (float)(hdx[1] * hdx[1]);
  sp.dz2 = (float)(hdx[2] * hdx[2]);
  sp.idx2 = 1.0f / sp.dx2;
  sp.idy2 = 1.0f / sp.dy2;
  sp.idz2 = 1.0f / sp.dz2;
  sp.jac = -2 / sp.dx2 - 2 / sp.dy2 - 2 / sp.dz2;
  sp.ijac = 1 / sp.jac;
  sp.pow2 = is_pow2(sp.dx2) && is_pow2(sp.dy2) && is_pow2(sp.dz2);
  return sp;
}

Spacing<double> spacing_of(const double *hdx) {
  Spacing<double> sp;
  sp.dx2 = hdx[0] * hdx[0];
  sp.dy2 = hdx[1] * hdx[1];
  sp.dz2 = hdx[2] * hdx[2];
  sp.idx2 = 1.0 / sp.dx2;
  sp.idy2 = 1.0 / sp.dy2;
  sp.idz2 = 1.0 / sp.dz2;
  sp.jac = -2 / sp.dx2 - 2 / sp.dy2 - 2 / sp.dz2;
  sp.ijac = 1 / sp.jac;
  sp.pow2 = is_pow2(sp.dx2) && is_pow2(sp.dy2) && is_pow2(sp.dz2);
  return sp;
}

// set_cmask_bdy / set_cmask_child of one grid (amrd/mg.c:362-488) as bytes
void cmask_bytes(const pmg_amr *A, int L, int gi, int which, std::vector<uint8_t> &m) {
  const HGrid &g = A->lev[L - 1].g[gi];
  const double *base = A->cbbox;
  m.assign(g.size(), 0);
  auto fill = [&](int is, int ie, int js, int je, int ks, int ke) {
    if (is >= 0 && js >= 0 && ks >= 0 && is <= ie && js <= je && ks <= ke && ie < g.n[0] && je < g.n[1] && ke < g.n[2])
      for (int k = ks; k <= ke; k++)
        for (int j = js; j <= je; j++)
          for (int i = is; i <= ie; i++) m[ix3(i, j, k, g.n[0], g.n[1])] = 1;
  };
  if (which == 0) {
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = 0;
      hi[d] = g.n[d] - 1;
      if ((g.bbox[2 * d] - base[2 * d]) > g.gdx[d] / 2 || A->periodic[d]) lo[d]++;
      if ((base[2 * d + 1] - g.bbox[2 * d + 1]) > g.gdx[d] / 2 || A->periodic[d]) hi[d]--;
    }
    fill(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
    return;
  }
  if (L >= A->nlev) return;
  for (const HGrid &cg : A->lev[L].g) {
    int s[3], e[3];
    for (int d = 0; d < 3; d++) {
      s[d] = (int)std::max((cg.bbox[2 * d] - g.bbox[2 * d]) / g.gdx[d] + 0.5, 0.0);
      e[d] = (int)std::min((cg.bbox[2 * d + 1] - g.bbox[2 * d]) / g.gdx[d] + 0.5, (double)(g.n[d] - 1));
      if (s[d] > 0 || A->periodic[d] || (g.bbox[2 * d] - base[2 * d]) > g.gdx[d] / 2) s[d]++;
      if (e[d] < (g.n[d] - 1) || A->periodic[d] || (base[2 * d + 1] - g.bbox[2 * d + 1]) > g.gdx[d] / 2) e[d]--;
    }
    fill(s[0], e[0], s[1], e[1], s[2], e[2]);
  }
}

// sweep tile shapes (TX, TY, NT); 1 and 2 are the defaults of single-grid
// and patch levels, the others are for the per-level autotuner
#define PMG_AMR_VARIANTS(X) \
  X(0, 128, 16, 1024) X(1, 64, 16, 512) X(2, 32, 16, 512) X(3, 32, 8, 256) X(4, 16, 16, 256)
constexpr int kAmrVariants = 5;

bool variant_txy(int v, int &tx, int &ty) {
  switch (v) {
#define PMG_TXY(V, TXV, TYV, NTV) case V: tx = TXV; ty = TYV; return true;
    PMG_AMR_VARIANTS(PMG_TXY)
#undef PMG_TXY
  }
  tx = 32;
  ty = 16;
  return false;
}

int build_tiles(pmg_amr *A, ALevel &Lv) {
  int tx, ty;
  variant_txy(Lv.variant, tx, ty);
  std::vector<BTile> st, pt;
  std::vector<int> ss{0}, ps{0};
  for (int gi = 0; gi < (int)Lv.g.size(); gi++) {
    const HGrid &g = Lv.g[gi];
    for (int z = 1; z <= g.n[2] - 2; z += Lv.zchunk)
      for (int x = 1; x <= g.n[0] - 2; x += tx)
        for (int y = 1; y <= g.n[1] - 2; y += ty) st.push_back(BTile{gi, x, y, z});
    ss.push_back((int)st.size());
    for (int z = 0; z < g.n[2]; z += PTZ)
      for (int y = 0; y < g.n[1]; y += PTY)
        for (int x = 0; x < g.n[0]; x += PTX) pt.push_back(BTile{gi, x, y, z});
    ps.push_back((int)pt.size());
  }
  Lv.nstiles = (int)st.size();
  Lv.nptiles = (int)pt.size();
  for (BTile **p : {&Lv.d_stiles, &Lv.d_ptiles}) if (*p) { (void)hipFree(*p); *p = nullptr; }
  for (int **p : {&Lv.d_sstart, &Lv.d_pstart}) if (*p) { (void)hipFree(*p); *p = nullptr; }
  PMG_HOST_CHECK(hipMalloc(&Lv.d_stiles, sizeof(BTile) * std::max<size_t>(1, st.size())));
  PMG_HOST_CHECK(hipMemcpy(Lv.d_stiles, st.data(), sizeof(BTile) * st.size(), hipMemcpyHostToDevice));
  PMG_HOST_CHECK(hipMalloc(&Lv.d_ptiles, sizeof(BTile) * pt.size()));
  PMG_HOST_CHECK(hipMemcpy(Lv.d_ptiles, pt.data(), sizeof(BTile) * pt.size(), hipMemcpyHostToDevice));
  PMG_HOST_CHECK(hipMalloc(&Lv.d_sstart, sizeof(int) * ss.size()));
  PMG_HOST_CHECK(hipMemcpy(Lv.d_sstart, ss.data(), sizeof(int) * ss.size(), hipMemcpyHostToDevice));
  PMG_HOST_CHECK(hipMalloc(&Lv.d_pstart, sizeof(int) * ps.size()));
  PMG_HOST_CHECK(hipMemcpy(Lv.d_pstart, ps.data(), sizeof(int) * ps.size(), hipMemcpyHostToDevice));
  return 0;
}

// ---------------------------------------------------------------------------
// level operations (enqueued on the solver stream)
// ---------------------------------------------------------------------------
double *field_pool(ALevel &Lv, int which) {
  switch (which) {
    case PMG_F_U: return Lv.u;
    case PMG_F_RHO: return Lv.rho;
    case PMG_F_RHS: return Lv.rhs;
    case PMG_F_RES: return Lv.res;
    case PMG_F_LOP: return Lv.lop;
    case PMG_F_A: return Lv.a;
    case PMG_F_C: return Lv.c;
  }
  if (which >= PMG_F_EXTRA && which - PMG_F_EXTRA < (int)Lv.extra.size()) return Lv.extra[which - PMG_F_EXTRA];
  return nullptr;
}

// the level's pools of one precision
template <typename T>
struct LvPools {
  T *u, *ua, *rho, *rhs, *res, *lop, *a, *c;
};
template <typename T>
LvPools<T> pools(ALevel &Lv);
template <>
LvPools<double> pools<double>(ALevel &Lv) {
  return {Lv.u, Lv.ua, Lv.rho, Lv.rhs, Lv.res, Lv.lop, Lv.a, Lv.c};
}
template <>
LvPools<float> pools<float>(ALevel &Lv) {
  return {Lv.f32[PMG_F_U], Lv.f32[7], Lv.f32[PMG_F_RHO], Lv.f32[PMG_F_RHS], Lv.f32[PMG_F_RES], Lv.f32[PMG_F_LOP],
          Lv.f32[PMG_F_A], Lv.f32[PMG_F_C]};
}
template <typename T>
const Spacing<T> &spacing_t(const ALevel &Lv);
template <>
const Spacing<double> &spacing_t<double>(const ALevel &Lv) { return Lv.sp; }
template <>
const Spacing<float> &spacing_t<float>(const ALevel &Lv) { return Lv.sp32; }

template <typename T>
T *field_pool_t(ALevel &Lv, int which) {
  if constexpr (std::is_same<T, double>::value) {
    return field_pool(Lv, which);
  } else {
    return (which >= PMG_F_U && which <= PMG_F_C) ? Lv.f32[which] : nullptr;
  }
}

bool valid_field(const pmg_amr *A, int which) {
  return (which >= 0 && which <= PMG_F_C) || (which >= PMG_F_EXTRA && which < PMG_F_EXTRA + A->nextra);
}

// one transfer plan with up to kMaxFields fields; kind 0 copy, 1-3 restrict,
// 4 interp (second order), 5 interp (fourth order)
template <typename T = double>
int run_plan(pmg_amr *A, const XPlan &P, int Ls, int Ld, int kind, const int *which, int nf) {
  if (P.nblocks == 0 || nf == 0) return 0;
  ALevel &S = A->lev[Ls - 1], &D = A->lev[Ld - 1];
  XFields<T> F;
  F.nf = 0;
  for (int f = 0; f < nf; f++) {
    T *s = field_pool_t<T>(S, which[f]), *d = field_pool_t<T>(D, which[f]);
    if (!s || !d) continue;
    F.src[F.nf] = s;
    F.dst[F.nf] = d;
    F.alt[F.nf] = which[f] == PMG_F_U ? pools<T>(D).ua : nullptr;
    F.nf++;
  }
  if (F.nf == 0) return 0;
  const dim3 grid(P.nblocks), blk(kXT);
  switch (kind) {
    case 0: hipLaunchKernelGGL(k_xfer_copy<T>, grid, blk, 0, A->stream, F, S.d_geo, D.d_geo, P.d_jobs, P.d_blks); break;
    case 1: hipLaunchKernelGGL((k_xfer_restrict<T, 1>), grid, blk, 0, A->stream, F, S.d_geo, D.d_geo, P.d_jobs, P.d_blks); break;
    case 2: hipLaunchKernelGGL((k_xfer_restrict<T, 2>), grid, blk, 0, A->stream, F, S.d_geo, D.d_geo, P.d_jobs, P.d_blks); break;
    case 3: hipLaunchKernelGGL((k_xfer_restrict<T, 3>), grid, blk, 0, A->stream, F, S.d_geo, D.d_geo, P.d_jobs, P.d_blks); break;
    case 4: hipLaunchKernelGGL(k_xfer_interp<T>, grid, blk, 0, A->stream, F, S.d_geo, D.d_geo, P.d_jobs, P.d_blks); break;
    case 5: hipLaunchKernelGGL(k_xfer_interp4<T>, grid, blk, 0, A->stream, F, S.d_geo, D.d_geo, P.d_jobs, P.d_blks); break;
    default: return -1;
  }
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

template <typename T = double>
int do_sync(pmg_amr *A, int L, int bdy, int which) {
  ALevel &Lv = A->lev[L - 1];
  return run_plan<T>(A, bdy ? Lv.sync1 : Lv.sync0, L, L, 0, &which, 1);
}

// PAMR_inject(Lf): u and u_res; the sources only into pure MG levels
// (TF_INJECT vs TF_INJECT_TO_MG_LEV, src/pamr.c:677-681,1886-1897)
template <typename T = double>
int do_inject_cycle(pmg_amr *A, int Lf) {
  ALevel &C = A->lev[Lf - 2];
  int ops[kMaxFields], ws[kMaxFields], n = 0;
  ops[n] = A->rop_u; ws[n++] = PMG_F_U;
  ops[n] = A->rop_res; ws[n++] = PMG_F_RES;
  if (!C.in_amrh) {
    ops[n] = A->rop_src; ws[n++] = PMG_F_RHO;
    if (A->op == 1) { ops[n] = A->rop_src; ws[n++] = PMG_F_A; ops[n] = A->rop_src; ws[n++] = PMG_F_C; }
  }
  for (int op = 1; op <= 3; op++) {
    int sel[kMaxFields], m = 0;
    for (int f = 0; f < n; f++) if (ops[f] == op) sel[m++] = ws[f];
    int rc = run_plan<T>(A, C.inject, Lf, Lf - 1, op, sel, m);
    if (rc) return rc;
  }
  return 0;
}

template <typename T = double>
int clear_mg_vars(pmg_amr *A, int L) {
  ALevel &Lv = A->lev[L - 1];
  const LvPools<T> P = pools<T>(Lv);
  if (!Lv.pool) return 0;
  if (!P.res || !P.rhs || !P.lop) return -1;
  hipLaunchKernelGGL(k_clear3<T>, dim3(nblk(Lv.pool)), dim3(256), 0, A->stream, P.res, P.rhs, P.lop, Lv.pool);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

template <typename T, int OP, bool P2, bool RHS0>
int launch_sweep_batch(pmg_amr *A, ALevel &Lv, const SweepBatchP<T> &B) {
  const dim3 grid(Lv.nstiles);
  switch (Lv.variant) {
#define PMG_LSB(V, TXV, TYV, NTV) \
  case V: hipLaunchKernelGGL((k_sweep_batch<T, OP, P2, RHS0, TXV, TYV, NTV>), grid, dim3(NTV), 0, A->stream, B); break;
    PMG_AMR_VARIANTS(PMG_LSB)
#undef PMG_LSB
    default: return -1;
  }
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

// n batched sweeps on level L, each followed by PAMR_sync(L, 0 then 1)
// (amrd/mg.c:135-141,161-167); sync writes both ping-pong buffers
template <typename T = double>
int sweeps(pmg_amr *A, int L, int n, bool do_sync_after) {
  ALevel &Lv = A->lev[L - 1];
  const LvPools<T> Q = pools<T>(Lv);
  if (n <= 0 || Lv.nstiles == 0) return 0;
  if (Lv.nstiles > A->ws.cap) return -3;
  SweepBatchP<T> B;
  B.rho = Q.rho; B.rhs = Q.rhs; B.a = Q.a; B.c = Q.c;
  B.grids = Lv.d_geo;
  B.tiles = Lv.d_stiles;
  B.zchunk = Lv.zchunk;
  B.sp = spacing_t<T>(Lv);
  B.ws = A->ws;
  B.slot = AS_RELAX + L - 1;
  const bool rhs0 = L == A->nlev, p2 = B.sp.pow2 != 0;
  for (int s = 0; s < n; s++) {
    T *cur = (s % 2 == 0) ? Q.u : Q.ua, *nxt = (s % 2 == 0) ? Q.ua : Q.u;
    B.uin = cur;
    B.uout = nxt;
    int rc;
#define PMG_SB(OPV) (p2 ? (rhs0 ? launch_sweep_batch<T, OPV, true, true>(A, Lv, B) : launch_sweep_batch<T, OPV, true, false>(A, Lv, B)) \
                        : (rhs0 ? launch_sweep_batch<T, OPV, false, true>(A, Lv, B) : launch_sweep_batch<T, OPV, false, false>(A, Lv, B)))
    rc = A->op == 0 ? PMG_SB(0) : PMG_SB(1);
#undef PMG_SB
    if (rc) return rc;
    if (do_sync_after) {
      const XPlan &P = s == 0 ? Lv.sync0 : Lv.sync1;
      if (P.nblocks) {
        XFields<T> F;
        F.nf = 1;
        F.src[0] = nxt;
        F.dst[0] = nxt;
        F.alt[0] = cur;
        hipLaunchKernelGGL(k_xfer_copy<T>, dim3(P.nblocks), dim3(kXT), 0, A->stream, F, Lv.d_geo, Lv.d_geo,
                           P.d_jobs, P.d_blks);
        PMG_HOST_CHECK(hipGetLastError());
      }
    }
  }
  if (n % 2 == 1) {
    hipLaunchKernelGGL(k_copy<T>, dim3(nblk(Lv.pool)), dim3(256), 0, A->stream, Q.u, Q.ua, Lv.pool);
    PMG_HOST_CHECK(hipGetLastError());
  }
  return 0;
}

template <typename T = double>
int residual(pmg_amr *A, int L, int slot) {
  ALevel &Lv = A->lev[L - 1];
  const LvPools<T> Q = pools<T>(Lv);
  if (Lv.nptiles > A->ws.cap) return -3;
  PtsBatchP<T> B{};
  B.u = Q.u; B.rho = Q.rho; B.rhs = Q.rhs; B.a = Q.a; B.c = Q.c;
  B.res = Q.res;
  B.grids = Lv.d_geo;
  B.tiles = Lv.d_ptiles;
  B.sp = spacing_t<T>(Lv);
  B.ws = A->ws;
  B.slot = slot;
  const bool rhs0 = L == A->nlev, p2 = B.sp.pow2 != 0;
  const dim3 grid(Lv.nptiles), blk(PTX * PTY);
#define PMG_RB(OPV, P2V, R0) hipLaunchKernelGGL((k_resid_batch<T, OPV, P2V, R0>), grid, blk, 0, A->stream, B)
  if (A->op == 0) {
    if (p2) { if (rhs0) PMG_RB(0, true, true); else PMG_RB(0, true, false); }
    else { if (rhs0) PMG_RB(0, false, true); else PMG_RB(0, false, false); }
  } else {
    if (p2) { if (rhs0) PMG_RB(1, true, true); else PMG_RB(1, true, false); }
    else { if (rhs0) PMG_RB(1, false, true); else PMG_RB(1, false, false); }
  }
#undef PMG_RB
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

template <typename T = double>
int lop_rhs(pmg_amr *A, int L) {
  ALevel &Lv = A->lev[L - 1];
  const LvPools<T> Q = pools<T>(Lv);
  PtsBatchP<T> B{};
  B.u = Q.u; B.rho = Q.rho; B.res = Q.res; B.a = Q.a; B.c = Q.c;
  B.lop = Q.lop;
  B.rhs_out = Q.rhs;
  B.child = Lv.child;
  B.grids = Lv.d_geo;
  B.tiles = Lv.d_ptiles;
  B.sp = spacing_t<T>(Lv);
  B.w0_r = (T)A->w0_r;
  const bool p2 = B.sp.pow2 != 0;
  const dim3 grid(Lv.nptiles), blk(PTX * PTY);
#define PMG_LB(OPV, P2V) hipLaunchKernelGGL((k_lop_rhs_batch<T, OPV, P2V>), grid, blk, 0, A->stream, B)
  if (A->op == 0) { if (p2) PMG_LB(0, true); else PMG_LB(0, false); }
  else { if (p2) PMG_LB(1, true); else PMG_LB(1, false); }
#undef PMG_LB
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

// compute_cgc (amrd/mg.c:323-355): res = (u - lop)/w0_i on L-1, PAMR_interp,
// u += res on L (whole grids, both ping-pong buffers)
template <typename T = double>
int cgc(pmg_amr *A, int L) {
  ALevel &C = A->lev[L - 2], &F = A->lev[L - 1];
  const LvPools<T> QC = pools<T>(C), QF = pools<T>(F);
  if (A->iord != 2 && A->iord != 4) return -4;
  if (F.interp_once) {
    const XPlan &P = F.interp;
    if (P.nblocks == 0) return 0;
    if (A->iord == 4)
      hipLaunchKernelGGL((k_xfer_cgc<T, 4>), dim3(P.nblocks), dim3(kXT), 0, A->stream, QC.u, QC.lop, (T)A->w0_i,
                         QF.u, QF.ua, C.d_geo, F.d_geo, P.d_jobs, P.d_blks);
    else
      hipLaunchKernelGGL((k_xfer_cgc<T, 2>), dim3(P.nblocks), dim3(kXT), 0, A->stream, QC.u, QC.lop, (T)A->w0_i,
                         QF.u, QF.ua, C.d_geo, F.d_geo, P.d_jobs, P.d_blks);
    PMG_HOST_CHECK(hipGetLastError());
    return 0;
  }
  hipLaunchKernelGGL(k_cgc_diff<T>, dim3(nblk(C.pool)), dim3(256), 0, A->stream, QC.res, QC.u, QC.lop, (T)A->w0_i,
                     C.pool);
  PMG_HOST_CHECK(hipGetLastError());
  const int w = PMG_F_RES;
  int rc = run_plan<T>(A, F.interp, L - 1, L, A->iord == 4 ? 5 : 4, &w, 1);
  if (rc) return rc;
  hipLaunchKernelGGL(k_add_alt<T>, dim3(nblk(F.pool)), dim3(256), 0, A->stream, QF.u, QF.ua, QF.res, F.pool);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

template <typename T = double>
int coarse_solve(pmg_amr *A) {
  ALevel &Lv = A->lev[0];
  const LvPools<T> Q = pools<T>(Lv);
  const HGrid &g = Lv.g[0];
  CoarseP<T> P;
  P.u = Q.u; P.rho = Q.rho; P.rhs = Q.rhs; P.a = Q.a; P.c = Q.c; P.chr = nullptr;
  P.nx = g.n[0]; P.ny = g.n[1]; P.nz = g.n[2];
  P.bx0 = P.by0 = P.bz0 = 1;
  P.bx1 = g.n[0] - 2; P.by1 = g.n[1] - 2; P.bz1 = g.n[2] - 2;
  P.sp = spacing_t<T>(Lv);
  P.npre = A->pre; P.npst = A->pst; P.max_citer = A->max_citer; P.crtol = A->crtol;
  P.norm_pre = A->d_cnorm; P.norm_post = A->d_cnorm + 1; P.citer = A->d_citer;
  P.snap = reinterpret_cast<T *>(A->d_snap);
  const int narr = (A->op == 1) ? 5 : 3;
  const size_t lds = g.size() * narr * sizeof(T);
  const bool use_lds = lds <= 60 * 1024, p2 = P.sp.pow2 != 0;
  constexpr int NT = 256;
#define PMG_CS(OPV, P2V, LDSV) hipLaunchKernelGGL((k_coarse_solve<T, OPV, false, P2V, LDSV, NT>), dim3(1), dim3(NT), LDSV ? lds : 0, A->stream, P)
#define PMG_CS2(OPV) \
  if (p2) { if (use_lds) PMG_CS(OPV, true, true); else PMG_CS(OPV, true, false); } \
  else { if (use_lds) PMG_CS(OPV, false, true); else PMG_CS(OPV, false, false); }
  if (A->op == 0) { PMG_CS2(0) } else { PMG_CS2(1) }
#undef PMG_CS2
#undef PMG_CS
  PMG_HOST_CHECK(hipGetLastError());
  // u and its ping-pong twin agree after the in-place coarse solve
  hipLaunchKernelGGL(k_copy<T>, dim3(nblk(Lv.pool)), dim3(256), 0, A->stream, Q.ua, Q.u, Lv.pool);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int fold(pmg_amr *A) {
  if (!A->d_fold) return -1;  // a failed re-tiling left no fold table
  hipLaunchKernelGGL(k_fold_max, dim3(A->nfold), dim3(256), 0, A->stream, A->ws, A->d_fold, A->d_norms);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

// amrd/mg.c:84-201; the FP32 instantiation is the inner cycle of mixed
// precision (no final residual: the FP64 defect replaces it)
template <typename T = double>
int enqueue_vcycle_t(pmg_amr *A) {
  const int Lf = A->nlev;
  int rc;
  if ((rc = clear_mg_vars<T>(A, Lf))) return rc;
  for (int L = Lf; L >= 1; L--) {
    if (L < Lf) {
      if ((rc = residual<T>(A, L + 1, AS_RESID + L))) return rc;
      if ((rc = do_sync<T>(A, L + 1, 0, PMG_F_RES))) return rc;  // set_res_sync + PAMR_sync
      if ((rc = clear_mg_vars<T>(A, L))) return rc;
      if ((rc = do_inject_cycle<T>(A, L + 1))) return rc;
      if ((rc = do_sync<T>(A, L, 0, PMG_F_U))) return rc;
      if ((rc = lop_rhs<T>(A, L))) return rc;
    }
    if (L == 1) {
      if ((rc = coarse_solve<T>(A))) return rc;
    } else {
      if ((rc = sweeps<T>(A, L, A->pre, true))) return rc;
    }
  }
  for (int L = 2; L <= Lf; L++) {
    if ((rc = cgc<T>(A, L))) return rc;
    if ((rc = sweeps<T>(A, L, A->pst, true))) return rc;
  }
  if constexpr (std::is_same<T, float>::value) return 0;
  if ((rc = residual<T>(A, Lf, AS_FINAL))) return rc;
  return fold(A);
}

int enqueue_vcycle(pmg_amr *A) { return enqueue_vcycle_t<double>(A); }

// FP32 shadows of every level (allocated once; the coefficients converted
// from FP64 at every call, so that later uploads of a and c are seen)
int ensure_fp32(pmg_amr *A) {
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  for (ALevel &Lv : A->lev) {
    if (!A->have_fp32)
      for (int w = 0; w < 8; w++) {
        if ((w == PMG_F_A || w == PMG_F_C) && A->op == 0) continue;
        PMG_HOST_CHECK(hipMalloc(&Lv.f32[w], Lv.pool * sizeof(float)));
        PMG_HOST_CHECK(hipMemset(Lv.f32[w], 0, Lv.pool * sizeof(float)));
      }
    if (A->op == 1) {
      hipLaunchKernelGGL(k_to_f32, dim3(nblk(Lv.pool)), dim3(256), 0, A->stream, Lv.f32[PMG_F_A], Lv.a, Lv.pool);
      hipLaunchKernelGGL(k_to_f32, dim3(nblk(Lv.pool)), dim3(256), 0, A->stream, Lv.f32[PMG_F_C], Lv.c, Lv.pool);
      PMG_HOST_CHECK(hipGetLastError());
    }
  }
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  A->have_fp32 = true;
  return 0;
}

// u of every coarser AMR level replaced, where a finer level covers it, by
// the restriction of the finer u (top down), then synchronised: the FAS cycle
// does this at the start of each level of its V (amrd/mg.c:112-118), so that
// the coarse values under a finer level -- those at the finer level's
// boundary included, whose coarse corrections become the finer boundary's --
// are the finer solution's.  Without it the FP64 u of the coarse points at a
// finer boundary would keep their own history, and defect correction would
// converge to another fixed point than the FAS cycle's.
int inject_u_down(pmg_amr *A) {
  if (A->rop_u < 1 || A->rop_u > 3) return 0;
  const int w = PMG_F_U;
  for (int L = A->nlev - 1; L >= A->nbase; L--) {
    int rc;
    if ((rc = run_plan<double>(A, A->lev[L - 1].inject, L + 1, L, A->rop_u, &w, 1))) return rc;
    if ((rc = do_sync<double>(A, L, 0, PMG_F_U))) return rc;
  }
  return 0;
}

// the FP64 defect of every AMR level into the FP32 sources, after the
// injection of u above; the finest level's norm goes to the final slot, then
// the fold
int enqueue_defect(pmg_amr *A) {
  int rc;
  if ((rc = inject_u_down(A))) return rc;
  for (int L = A->nbase; L <= A->nlev; L++) {
    ALevel &Lv = A->lev[L - 1];
    if (Lv.nptiles > A->ws.cap) return -3;
    PtsBatchP<double> B{};
    B.u = Lv.u; B.rho = Lv.rho; B.a = Lv.a; B.c = Lv.c;
    B.grids = Lv.d_geo;
    B.tiles = Lv.d_ptiles;
    B.sp = Lv.sp;
    B.ws = A->ws;
    B.slot = L == A->nlev ? AS_FINAL : AS_RESID + L - 1;
    const bool p2 = Lv.sp.pow2 != 0;
    const dim3 grid(Lv.nptiles), blk(PTX * PTY);
    float *src = Lv.f32[PMG_F_RHO];
    if (A->op == 0) {
      if (p2) hipLaunchKernelGGL((k_defect32_batch<0, true>), grid, blk, 0, A->stream, B, src);
      else hipLaunchKernelGGL((k_defect32_batch<0, false>), grid, blk, 0, A->stream, B, src);
    } else {
      if (p2) hipLaunchKernelGGL((k_defect32_batch<1, true>), grid, blk, 0, A->stream, B, src);
      else hipLaunchKernelGGL((k_defect32_batch<1, false>), grid, blk, 0, A->stream, B, src);
    }
    PMG_HOST_CHECK(hipGetLastError());
  }
  return fold(A);
}

// one mixed-precision iteration: the FP32 V-cycle on A e = -r from e = 0,
// u += e on every AMR level, the new FP64 defect
int enqueue_mixed(pmg_amr *A) {
  int rc;
  for (ALevel &Lv : A->lev) {
    hipLaunchKernelGGL(k_fill0<float>, dim3(nblk(Lv.pool)), dim3(256), 0, A->stream, Lv.f32[PMG_F_U], Lv.f32[7],
                       Lv.pool);
    PMG_HOST_CHECK(hipGetLastError());
  }
  if ((rc = enqueue_vcycle_t<float>(A))) return rc;
  for (int L = A->nbase; L <= A->nlev; L++) {
    ALevel &Lv = A->lev[L - 1];
    hipLaunchKernelGGL(k_add32_alt, dim3(nblk(Lv.pool)), dim3(256), 0, A->stream, Lv.u, Lv.ua, Lv.f32[PMG_F_U],
                       Lv.pool);
    PMG_HOST_CHECK(hipGetLastError());
  }
  return enqueue_defect(A);
}

int enqueue_residual(pmg_amr *A) {
  int rc;
  const int Lf = A->nlev;
  if ((rc = residual(A, Lf, AS_FINAL))) return rc;
  return fold(A);
}

int capture_or_run(pmg_amr *A, hipGraphExec_t *cache, int (*enq)(pmg_amr *)) {
  if (!A->use_graph) return enq(A);
  if (!*cache) {
    PMG_HOST_CHECK(hipStreamBeginCapture(A->stream, hipStreamCaptureModeThreadLocal));
    int rc = enq(A);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(A->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) { std::fprintf(stderr, "pamr_amr: graph capture failed: %s\n", hipGetErrorString(e)); return -1; }
    PMG_HOST_CHECK(hipGraphInstantiate(cache, g, nullptr, nullptr, 0));
    PMG_HOST_CHECK(hipGraphDestroy(g));
  }
  PMG_HOST_CHECK(hipGraphLaunch(*cache, A->stream));
  return 0;
}

void drop_graphs(pmg_amr *A) {
  for (hipGraphExec_t *g : {&A->graph, &A->graph_res, &A->graph_mixed})
    if (*g) { (void)hipGraphExecDestroy(*g); *g = nullptr; }
}

void free_plan(XPlan &P) {
  if (P.d_jobs) (void)hipFree(P.d_jobs);
  if (P.d_blks) (void)hipFree(P.d_blks);
  P.d_jobs = nullptr;
  P.d_blks = nullptr;
}

int build_folds(pmg_amr *A) {
  std::vector<FoldSpec> fs(kAmrSlots);
  for (int s = 0; s < kAmrSlots; s++) fs[s] = FoldSpec{s, 0, nullptr};
  for (int L = 1; L <= A->nlev; L++) {
    ALevel &Lv = A->lev[L - 1];
    const int ng = (int)Lv.g.size();
    fs[AS_RELAX + L - 1] = FoldSpec{AS_RELAX + L - 1, L == 1 ? 0 : ng, Lv.d_sstart};
    fs[AS_RESID + L - 1] = FoldSpec{AS_RESID + L - 1, ng, Lv.d_pstart};
  }
  fs[AS_FINAL] = FoldSpec{AS_FINAL, (int)A->lev[A->nlev - 1].g.size(), A->lev[A->nlev - 1].d_pstart};
  A->nfold = kAmrSlots;
  PMG_HOST_CHECK(hipMalloc(&A->d_fold, sizeof(FoldSpec) * fs.size()));
  PMG_HOST_CHECK(hipMemcpy(A->d_fold, fs.data(), sizeof(FoldSpec) * fs.size(), hipMemcpyHostToDevice));
  return 0;
}

// per-level device state of a hierarchy whose level grids are set: pools
// (aliased from the parent on shared levels), grid tables, child masks, sweep
// tiles, transfer plans, norm workspace and fold table.  Returns 0 / -1.
int setup_levels(pmg_amr *A, int op) {
  if (A->own_stream && hipStreamCreateWithFlags(&A->stream, hipStreamNonBlocking) != hipSuccess) return -1;
  int cap = 1 << 14;
  for (int L = 1; L <= A->nlev; L++) {
    ALevel &Lv = A->lev[L - 1];
    // all grids of a level share one spacing (dyadic boxes); check it
    for (const HGrid &g : Lv.g)
      for (int d = 0; d < 3; d++)
        if (g.hdx[d] != Lv.g[0].hdx[d] || g.n[d] < 3) return -1;
    Lv.sp = spacing_of(Lv.g[0].hdx);
    Lv.sp32 = spacing_of32(Lv.g[0].hdx);
    size_t pool = 0;
    std::vector<GridGeo> geo;
    for (HGrid &g : Lv.g) {
      g.off = pool;
      geo.push_back(GridGeo{pool, g.n[0], g.n[1], g.n[2]});
      pool += (g.size() + 63) / 64 * 64;
    }
    Lv.pool = pool;
    if (Lv.shared) {
      const ALevel &P = A->parent->lev[Lv.plev - 1];
      if (P.pool != pool) return -1;
      Lv.u = P.u; Lv.ua = P.ua; Lv.rho = P.rho; Lv.rhs = P.rhs; Lv.res = P.res; Lv.lop = P.lop;
      Lv.a = P.a; Lv.c = P.c;
      Lv.extra = P.extra;
    } else {
      for (double **p : {&Lv.u, &Lv.ua, &Lv.rho, &Lv.rhs, &Lv.res, &Lv.lop}) {
        if (hipMalloc(p, pool * sizeof(double)) != hipSuccess) return -1;
        (void)hipMemset(*p, 0, pool * sizeof(double));
      }
      if (op == 1)
        for (double **p : {&Lv.a, &Lv.c}) {
          if (hipMalloc(p, pool * sizeof(double)) != hipSuccess) return -1;
          (void)hipMemset(*p, 0, pool * sizeof(double));
        }
      for (int f = 0; f < A->nextra; f++) {  // the parent's extra fields, zero on new levels
        double *p = nullptr;
        if (hipMalloc(&p, pool * sizeof(double)) != hipSuccess) return -1;
        (void)hipMemset(p, 0, pool * sizeof(double));
        Lv.extra.push_back(p);
      }
    }
    if (hipMalloc(&Lv.d_geo, sizeof(GridGeo) * geo.size()) != hipSuccess) return -1;
    (void)hipMemcpy(Lv.d_geo, geo.data(), sizeof(GridGeo) * geo.size(), hipMemcpyHostToDevice);
    // child mask (set_cmask_child) as bytes, fixed for the hierarchy
    std::vector<uint8_t> child(pool, 0), m;
    for (int gi = 0; gi < (int)Lv.g.size(); gi++) {
      cmask_bytes(A, L, gi, 1, m);
      std::copy(m.begin(), m.end(), child.begin() + Lv.g[gi].off);
    }
    if (hipMalloc(&Lv.child, pool) != hipSuccess) return -1;
    (void)hipMemcpy(Lv.child, child.data(), pool, hipMemcpyHostToDevice);
    // sweep tiles: single-grid levels use the wide tile, patch levels 32 x 16
    const HGrid &g0 = Lv.g[0];
    if (Lv.g.size() == 1) {
      Lv.variant = 1;
      Lv.zchunk = g0.n[2] >= 257 ? 32 : g0.n[2] >= 129 ? 16 : g0.n[2] >= 65 ? 8 : 4;
    } else {
      Lv.variant = 2;
      Lv.zchunk = 8;
    }
    if (build_tiles(A, Lv)) return -1;
    cap = std::max(cap, std::max(Lv.nstiles, Lv.nptiles) + 64);
  }
  // transfer plans
  for (int L = 1; L <= A->nlev; L++) {
    ALevel &Lv = A->lev[L - 1];
    std::vector<XJob> j0, j1, ji, jp;
    int rc = make_jobs(A, L, L, 1, 0, j0);
    if (!rc) rc = make_jobs(A, L, L, 2, 0, j1);
    if (!rc && L < A->nlev) rc = make_jobs(A, L + 1, L, 1, 0, ji);
    if (!rc && L > 1) rc = make_jobs(A, L - 1, L, 0, 1, jp);
    if (!rc) rc = upload_plan(A, Lv.sync0, j0);
    if (!rc) rc = upload_plan(A, Lv.sync1, j1);
    if (!rc) rc = upload_plan(A, Lv.inject, ji);
    if (!rc) rc = upload_plan(A, Lv.interp, jp);
    {  // a level with coarsest and other grids: the coarsest loop's sync among the coarsest
      int nco = 0;
      for (const HGrid &g : Lv.g) nco += g.coarsest ? 1 : 0;
      if (!rc && nco > 0 && nco < (int)Lv.g.size()) {
        std::vector<XJob> jc;
        rc = make_jobs(A, L, L, 2, 0, jc, true);
        if (!rc) rc = upload_plan(A, Lv.sync1c, jc);
      }
    }
    if (!rc && L > 1) {  // does the (resolved) interp plan write every point of every grid of L exactly once?
      std::vector<std::vector<uint8_t>> cnt(Lv.g.size());
      for (size_t gi = 0; gi < Lv.g.size(); gi++) cnt[gi].assign(Lv.g[gi].size(), 0);
      bool once = true;
      for (const XJob &J : Lv.interp.jobs)
        for (int k = J.lo[2]; k <= J.hi[2] && once; k++)
          for (int j = J.lo[1]; j <= J.hi[1] && once; j++)
            for (int i = J.lo[0]; i <= J.hi[0]; i++) {
              const HGrid &gg = Lv.g[J.dg];
              uint8_t &c = cnt[J.dg][((size_t)k * gg.n[1] + j) * gg.n[0] + i];
              if (c) { once = false; break; }
              c = 1;
            }
      for (size_t gi = 0; gi < cnt.size() && once; gi++)
        for (uint8_t c : cnt[gi])
          if (!c) { once = false; break; }
      Lv.interp_once = once;
    }
    if (rc) { std::fprintf(stderr, "pamr_amr: transfer plan of level %d failed (%d)\n", L, rc); return -1; }
  }
  A->ws.cap = cap;
  A->ws.nslots = kAmrSlots;
  if (hipMalloc(&A->ws.partials, sizeof(double) * 2 * (size_t)cap * kAmrSlots) != hipSuccess ||
      hipMalloc(&A->ws.slot_nb, sizeof(int) * kAmrSlots) != hipSuccess ||
      hipMalloc(&A->d_norms, sizeof(double) * kAmrSlots) != hipSuccess ||
      hipMalloc(&A->d_cnorm, sizeof(double) * 2) != hipSuccess || hipMalloc(&A->d_citer, sizeof(int) * 2) != hipSuccess ||
      hipMalloc(&A->d_snap, sizeof(double) * kSnap * std::max<size_t>(A->lev[0].pool, 1)) != hipSuccess ||
      hipHostMalloc(&A->h_pinned, sizeof(double) * (kAmrSlots + 4)) != hipSuccess)
    return -1;
  (void)hipMemset(A->ws.partials, 0, sizeof(double) * 2 * (size_t)cap * kAmrSlots);
  (void)hipMemset(A->d_norms, 0, sizeof(double) * kAmrSlots);
  (void)hipMemset(A->d_cnorm, 0, sizeof(double) * 2);
  (void)hipMemset(A->d_citer, 0, sizeof(int) * 2);
  if (build_folds(A)) return -1;
  if (hipDeviceSynchronize() != hipSuccess) return -1;
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

pmg_amr *pmg_amr_create(const int *base_shape, const double *bbox, int namr, const int *nbox, const double *boxes,
                        int min_mg_cwidth, int op) {
  if (!base_shape || !bbox || namr < 0 || (namr > 0 && (!nbox || !boxes)) || (op != 0 && op != 1)) return nullptr;
  pmg_amr *A = new pmg_amr();
  std::memcpy(A->cbbox, bbox, 6 * sizeof(double));
  A->op = op;
  double dx0[3];
  for (int d = 0; d < 3; d++) dx0[d] = (bbox[2 * d + 1] - bbox[2 * d]) / (base_shape[d] - 1);
  int ncoarse = 0;
  for (int cf = 2;; cf *= 2) {  // build_mgh below the base (src/gh.c:1275-1340)
    if (island_single(bbox, dx0, cf, min_mg_cwidth) == 0 || ncoarse + 1 + namr >= kMaxLev) break;
    ncoarse++;
  }
  A->nbase = ncoarse + 1;
  A->nlev = ncoarse + 1 + namr;
  A->lev.resize(A->nlev);
  std::memcpy(A->dx0, dx0, sizeof dx0);
  A->mcw = min_mg_cwidth;
  A->sgh.assign(1, std::vector<double>(bbox, bbox + 6));
  for (int l = 0, o = 0; l < namr; o += nbox[l], l++) {
    if (nbox[l] < 1) { delete A; return nullptr; }
    A->sgh.emplace_back(boxes + 6 * o, boxes + 6 * (o + nbox[l]));
  }
  for (int L = 1; L <= ncoarse; L++) {
    const int c = 1 << (ncoarse + 1 - L);
    const double ldx[3] = {dx0[0] * c, dx0[1] * c, dx0[2] * c};
    A->lev[L - 1].g = level_grids(1, bbox, ldx);
  }
  A->lev[A->nbase - 1].g = level_grids(1, bbox, dx0);
  A->lev[A->nbase - 1].in_amrh = true;
  int off = 0;
  for (int l = 0; l < namr; l++) {
    const double f = 1.0 / (double)(2 << l);
    const double ldx[3] = {dx0[0] * f, dx0[1] * f, dx0[2] * f};
    if (nbox[l] < 1) { delete A; return nullptr; }
    A->lev[A->nbase + l].g = level_grids(nbox[l], boxes + 6 * off, ldx);
    A->lev[A->nbase + l].in_amrh = true;
    off += nbox[l];
  }
  for (HGrid &g : A->lev[0].g) g.coarsest = true;
  if (A->lev[0].g.size() != 1) { delete A; return nullptr; }
  if (setup_levels(A, op)) { pmg_amr_destroy(A); return nullptr; }
  return A;
}

void pmg_amr_destroy(pmg_amr *A) {
  if (!A) return;
  if (A->stream) (void)hipStreamSynchronize(A->stream);
  drop_graphs(A);
  for (ALevel &Lv : A->lev) {
    if (!Lv.shared) {
      for (double **p : {&Lv.u, &Lv.ua, &Lv.rho, &Lv.rhs, &Lv.res, &Lv.lop, &Lv.a, &Lv.c})
        if (*p) (void)hipFree(*p);
      for (double *p : Lv.extra) if (p) (void)hipFree(p);
    }
    for (double *p : Lv.cm) if (p) (void)hipFree(p);
    if (Lv.child) (void)hipFree(Lv.child);
    for (float *p : Lv.f32) if (p) (void)hipFree(p);
    if (Lv.d_geo) (void)hipFree(Lv.d_geo);
    for (BTile *p : {Lv.d_stiles, Lv.d_ptiles}) if (p) (void)hipFree(p);
    for (int *p : {Lv.d_sstart, Lv.d_pstart}) if (p) (void)hipFree(p);
    for (XPlan *P : {&Lv.sync0, &Lv.sync1, &Lv.sync1c, &Lv.inject, &Lv.interp}) free_plan(*P);
  }
  if (A->xtmp) (void)hipFree(A->xtmp);
  if (A->ws.partials) (void)hipFree(A->ws.partials);
  if (A->ws.slot_nb) (void)hipFree(A->ws.slot_nb);
  for (double *p : {A->d_norms, A->d_cnorm, A->d_snap}) if (p) (void)hipFree(p);
  if (A->d_citer) (void)hipFree(A->d_citer);
  if (A->d_fold) (void)hipFree(A->d_fold);
  if (A->h_pinned) (void)hipHostFree(A->h_pinned);
  if (A->stream && A->own_stream) (void)hipStreamDestroy(A->stream);
  delete A;
}

pmg_amr *pmg_amr_create_mgh(pmg_amr *P, int amr_lev) {
  if (!P || P->parent || amr_lev < 1 || amr_lev > (int)P->sgh.size()) return nullptr;
  pmg_amr *A = new pmg_amr();
  std::memcpy(A->cbbox, P->cbbox, sizeof A->cbbox);
  std::memcpy(A->periodic, P->periodic, sizeof A->periodic);
  A->op = P->op;
  A->nextra = P->nextra;
  A->pre = P->pre; A->pst = P->pst; A->max_citer = P->max_citer;
  A->rop_u = P->rop_u; A->rop_res = P->rop_res; A->rop_src = P->rop_src; A->iord = P->iord;
  A->crtol = P->crtol; A->w0_r = P->w0_r; A->w0_i = P->w0_i;
  A->mcw = P->mcw;
  std::memcpy(A->dx0, P->dx0, sizeof A->dx0);
  A->parent = P;
  A->stream = P->stream;  // one stream for both hierarchies: transfers and hooks stay ordered
  A->own_stream = false;
  const std::vector<double> &bb = P->sgh[amr_lev - 1];
  const int num = (int)bb.size() / 6;
  double dx[3];
  for (int d = 0; d < 3; d++) dx[d] = P->dx0[d] / (double)(1 << (amr_lev - 1));
  const int Ls = P->nbase + amr_lev - 1;  // the parent's level holding AMR level amr_lev
  const int nsh = P->nlev - Ls + 1;
  // build_mgh below min_lev (src/gh.c:1281-1340): islands of the level's boxes
  std::vector<std::vector<HGrid>> coarse;  // finest first
  for (int cf = 2; (int)coarse.size() + nsh < kMaxLev; cf *= 2) {
    const std::vector<int> isl = island_numbers(bb.data(), num, dx, cf, P->mcw);
    std::vector<double> keep;
    std::vector<char> kc;
    for (int k = 0; k < num; k++)
      if (isl[k] != 0) {
        keep.insert(keep.end(), bb.begin() + 6 * k, bb.begin() + 6 * k + 6);
        kc.push_back(isl[k] < 0);
      }
    if (kc.empty()) break;
    const double ldx[3] = {dx[0] * cf, dx[1] * cf, dx[2] * cf};
    std::vector<int> perm;
    std::vector<HGrid> g = level_grids_perm((int)kc.size(), keep.data(), ldx, perm);
    for (size_t k = 0; k < g.size(); k++) g[k].coarsest = kc[perm[k]] != 0;  // fill_mgh_level (src/gh.c:1192)
    coarse.push_back(std::move(g));
  }
  const int nc = (int)coarse.size();
  A->nbase = nc + 1;
  A->nlev = nc + nsh;
  A->lev.resize(A->nlev);
  for (int L = 1; L <= nc; L++) A->lev[L - 1].g = coarse[nc - L];
  for (int j = 0; j < nsh; j++) {  // amr_to_mgh (src/gh.c:1097-1165)
    ALevel &Lv = A->lev[nc + j];
    Lv.g = P->lev[Ls - 1 + j].g;
    for (HGrid &g : Lv.g) g.coarsest = false;
    Lv.in_amrh = true;
    Lv.shared = true;
    Lv.plev = Ls + j;
  }
  for (HGrid &g : A->lev[0].g) g.coarsest = true;  // src/gh.c:1352-1354
  if (setup_levels(A, A->op)) {
    pmg_amr_destroy(A);
    return nullptr;
  }
  return A;
}

int pmg_amr_grid_coarsest(pmg_amr *A, int L, int g) {
  if (!A || L < 1 || L > A->nlev || g < 0 || g >= (int)A->lev[L - 1].g.size()) return -1;
  return A->lev[L - 1].g[g].coarsest ? 1 : 0;
}

int pmg_amr_nlev(pmg_amr *A) { return A ? A->nlev : -1; }
int pmg_amr_base_level(pmg_amr *A) { return A ? A->nbase : -1; }
int pmg_amr_ngrids(pmg_amr *A, int L) { return (A && L >= 1 && L <= A->nlev) ? (int)A->lev[L - 1].g.size() : -1; }

static bool valid_grid(pmg_amr *A, int L, int g) {
  return A && L >= 1 && L <= A->nlev && g >= 0 && g < (int)A->lev[L - 1].g.size();
}

int pmg_amr_grid_info(pmg_amr *A, int L, int g, int *shape, double *bbox) {
  if (!valid_grid(A, L, g)) return -1;
  const HGrid &gr = A->lev[L - 1].g[g];
  if (shape) std::memcpy(shape, gr.n, 3 * sizeof(int));
  if (bbox) std::memcpy(bbox, gr.bbox, 6 * sizeof(double));
  return 0;
}

double *pmg_amr_field(pmg_amr *A, int L, int g, int which) {
  if (!valid_grid(A, L, g)) return nullptr;
  double *p = field_pool(A->lev[L - 1], which);
  return p ? p + A->lev[L - 1].g[g].off : nullptr;
}

int pmg_amr_commit_u(pmg_amr *A) {
  if (!A) return -1;
  for (ALevel &Lv : A->lev)
    hipLaunchKernelGGL(k_copy<double>, dim3(nblk(Lv.pool)), dim3(256), 0, A->stream, Lv.ua, Lv.u, Lv.pool);
  PMG_HOST_CHECK(hipGetLastError());
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  return 0;
}

int pmg_amr_upload(pmg_amr *A, int L, int g, int which, const double *host) {
  double *p = pmg_amr_field(A, L, g, which);
  if (!p || !host) return -1;
  const HGrid &gr = A->lev[L - 1].g[g];
  PMG_HOST_CHECK(hipMemcpyAsync(p, host, gr.size() * sizeof(double), hipMemcpyHostToDevice, A->stream));
  if (which == PMG_F_U)
    PMG_HOST_CHECK(hipMemcpyAsync(A->lev[L - 1].ua + gr.off, host, gr.size() * sizeof(double), hipMemcpyHostToDevice,
                                  A->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  return 0;
}

int pmg_amr_download(pmg_amr *A, int L, int g, int which, double *host) {
  double *p = pmg_amr_field(A, L, g, which);
  if (!p || !host) return -1;
  const HGrid &gr = A->lev[L - 1].g[g];
  PMG_HOST_CHECK(hipMemcpyAsync(host, p, gr.size() * sizeof(double), hipMemcpyDeviceToHost, A->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  return 0;
}

int pmg_amr_set_params(pmg_amr *A, int pre, int pst, int max_citer, double crtol, double w0_r, double w0_i, int rop_u,
                       int rop_res, int rop_src, int interp_ord) {
  if (!A || pre < 0 || pst < 0 || max_citer < 0) return -1;
  for (int r : {rop_u, rop_res, rop_src})
    if (r < 0 || r > 3) return -1;
  if (interp_ord != 2 && interp_ord != 4) return -2;
  A->pre = pre; A->pst = pst; A->max_citer = max_citer; A->crtol = crtol;
  A->w0_r = w0_r; A->w0_i = w0_i; A->rop_u = rop_u; A->rop_res = rop_res; A->rop_src = rop_src;
  A->iord = interp_ord;
  drop_graphs(A);
  return 0;
}

int pmg_amr_init(pmg_amr *A) {
  if (!A) return -1;
  for (int L = A->nbase; L > 1; L--) {  // TF_MGH_INIT: u and the sources (in_amrh with mg_inject)
    int sel[kMaxFields], ops[kMaxFields], n = 0;
    ops[n] = A->rop_u; sel[n++] = PMG_F_U;
    ops[n] = A->rop_src; sel[n++] = PMG_F_RHO;
    if (A->op == 1) { ops[n] = A->rop_src; sel[n++] = PMG_F_A; ops[n] = A->rop_src; sel[n++] = PMG_F_C; }
    for (int op = 1; op <= 3; op++) {
      int s2[kMaxFields], m = 0;
      for (int f = 0; f < n; f++) if (ops[f] == op) s2[m++] = sel[f];
      int rc = run_plan(A, A->lev[L - 2].inject, L, L - 1, op, s2, m);
      if (rc) return rc;
    }
  }
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  return 0;
}

static int read_norms(pmg_amr *A, double *res_out, double *lev_norms, int *citer) {
  PMG_HOST_CHECK(hipMemcpyAsync(A->h_pinned, A->d_norms, sizeof(double) * kAmrSlots, hipMemcpyDeviceToHost, A->stream));
  PMG_HOST_CHECK(hipMemcpyAsync(A->h_pinned + kAmrSlots, A->d_cnorm, sizeof(double) * 2, hipMemcpyDeviceToHost, A->stream));
  int ci[2];
  PMG_HOST_CHECK(hipMemcpyAsync(ci, A->d_citer, sizeof(int) * 2, hipMemcpyDeviceToHost, A->stream));
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  if (res_out) *res_out = A->h_pinned[AS_FINAL];
  if (lev_norms) {
    lev_norms[0] = A->h_pinned[kAmrSlots + 1];
    for (int L = 2; L <= A->nlev; L++) lev_norms[L - 1] = A->h_pinned[AS_RELAX + L - 1];
    lev_norms[A->nlev - 1] = A->h_pinned[AS_FINAL];
  }
  if (citer) *citer = ci[0];
  return 0;
}

int pmg_amr_vcycle(pmg_amr *A, double *res_out, double *lev_norms, int *citer) {
  if (!A) return -1;
  if (A->parent) {  // the built-in cycle needs one coarsest grid; MG hierarchies from a refined level run the hooks
    std::fprintf(stderr, "pamr_amr: the built-in V-cycle runs on hierarchies from level 1 only\n");
    return -1;
  }
  int rc = capture_or_run(A, &A->graph, enqueue_vcycle);
  if (rc) return rc;
  return read_norms(A, res_out, lev_norms, citer);
}

int pmg_amr_solve(pmg_amr *A, int min_iter, int max_iter, double tol, double *norms, int *iters) {
  if (!A) return -1;
  if (A->parent) {  // the built-in cycle needs one coarsest grid; MG hierarchies from a refined level run the hooks
    std::fprintf(stderr, "pamr_amr: the built-in V-cycle runs on hierarchies from level 1 only\n");
    return -1;
  }
  int it = 0;
  double res = tol + 1;
  while ((it < min_iter || res > tol) && it < max_iter) {
    int rc = pmg_amr_vcycle(A, &res, nullptr, nullptr);
    if (rc) return rc;
    if (norms) norms[it] = res;
    it++;
  }
  if (iters) *iters = it;
  return 0;
}

int pmg_amr_residual(pmg_amr *A, double *norm_out) {
  if (!A) return -1;
  if (A->parent) {  // the built-in cycle needs one coarsest grid; MG hierarchies from a refined level run the hooks
    std::fprintf(stderr, "pamr_amr: the built-in V-cycle runs on hierarchies from level 1 only\n");
    return -1;
  }
  int rc = capture_or_run(A, &A->graph_res, enqueue_residual);
  if (rc) return rc;
  return read_norms(A, norm_out, nullptr, nullptr);
}

int pmg_amr_solve_rel(pmg_amr *A, int mode, double rtol, int max_iter, double *hist, int *iters, double *r0_out) {
  if (!A || max_iter < 0 || (mode != 0 && mode != 1)) return -1;
  if (A->parent) {
    std::fprintf(stderr, "pamr_amr: the built-in V-cycle runs on hierarchies from level 1 only\n");
    return -1;
  }
  int rc, it = 0;
  double r;
  if ((rc = pmg_amr_residual(A, &r))) return rc;
  const double rref = (r0_out && *r0_out > 0) ? *r0_out : r;
  if (r0_out) *r0_out = rref;
  if (mode == 1 && (rc = ensure_fp32(A))) return rc;
  bool defect_ready = false;  // the FP32 sources hold the defect of the current u
  while (it < max_iter && r > rtol * rref) {
    if (mode == 0 || r > A->mixed_switch * rref) {
      if ((rc = pmg_amr_vcycle(A, &r, nullptr, nullptr))) return rc;
      defect_ready = false;
    } else {
      if (!defect_ready) {
        if ((rc = enqueue_defect(A))) return rc;
        defect_ready = true;
      }
      if ((rc = capture_or_run(A, &A->graph_mixed, enqueue_mixed))) return rc;
      if ((rc = read_norms(A, &r, nullptr, nullptr))) return rc;
    }
    if (hist) hist[it] = r;
    it++;
  }
  if (iters) *iters = it;
  return 0;
}

int pmg_amr_set_mixed_switch(pmg_amr *A, double frac) {
  if (!A || !(frac >= 0.0) || frac > 1.0) return -1;
  A->mixed_switch = frac;
  return 0;
}

int pmg_amr_sync(pmg_amr *A, int L, int bdy, int which) {
  if (!A || L < 1 || L > A->nlev || !valid_field(A, which) || (bdy != 0 && bdy != 1)) return -1;
  int rc = do_sync(A, L, bdy, which);
  if (rc) return rc;
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  return 0;
}

int pmg_amr_inject(pmg_amr *A, int Lf, int which, int op) {
  if (!A || Lf < 2 || Lf > A->nlev || !valid_field(A, which) || op < 1 || op > 3) return -1;
  int rc = run_plan(A, A->lev[Lf - 2].inject, Lf, Lf - 1, op, &which, 1);
  if (rc) return rc;
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  return 0;
}

int pmg_amr_interp(pmg_amr *A, int Lc, int which) { return pmg_amr_interp_ord(A, Lc, which, 2); }

int pmg_amr_interp_ord(pmg_amr *A, int Lc, int which, int ord) {
  if (!A || Lc < 1 || Lc >= A->nlev || !valid_field(A, which)) return -1;
  if (ord != 2 && ord != 4) return -2;
  int rc = run_plan(A, A->lev[Lc].interp, Lc, Lc + 1, ord == 4 ? 5 : 4, &which, 1);
  if (rc) return rc;
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  return 0;
}

int pmg_amr_cmask(pmg_amr *A, int L, int g, int which, double *host_out) {
  if (!valid_grid(A, L, g) || !host_out || (which != 0 && which != 1)) return -1;
  std::vector<uint8_t> m;
  cmask_bytes(A, L, g, which, m);
  for (size_t q = 0; q < m.size(); q++) host_out[q] = m[q] ? 1.0 : 0.0;
  return 0;
}

int pmg_amr_plan_stats(pmg_amr *A, int L, int *out) {
  if (!A || L < 1 || L > A->nlev || !out) return -1;
  ALevel &Lv = A->lev[L - 1];
  out[0] = (int)Lv.sync0.jobs.size();
  out[1] = (int)Lv.sync1.jobs.size();
  out[2] = (int)Lv.inject.jobs.size();
  out[3] = (int)Lv.interp.jobs.size();
  return 0;
}

int pmg_amr_time_sweep(pmg_amr *A, int L, int nrep, double *ms_avg) {
  if (!A || L < 2 || L > A->nlev || nrep < 1 || !ms_avg) return -1;
  EventPair ev;
  PMG_HOST_CHECK(ev.create());
  hipEvent_t e0 = ev.e0, e1 = ev.e1;
  int rc = sweeps(A, L, 2, false);  // warm up (even count: u returns to its buffer)
  if (rc) return rc;
  PMG_HOST_CHECK(hipEventRecord(e0, A->stream));
  rc = sweeps(A, L, 2 * nrep, false);
  if (rc) return rc;
  PMG_HOST_CHECK(hipEventRecord(e1, A->stream));
  PMG_HOST_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  PMG_HOST_CHECK(hipEventElapsedTime(&ms, e0, e1));
  *ms_avg = ms / (2.0 * nrep);
  return 0;
}

int pmg_amr_set_graph(pmg_amr *A, int on) {
  if (!A) return -1;
  A->use_graph = on ? 1 : 0;
  drop_graphs(A);
  return 0;
}

void *pmg_amr_stream(pmg_amr *A) { return A ? (void *)A->stream : nullptr; }

int pmg_amr_add_field(pmg_amr *A) {
  if (!A || A->nextra >= PMG_MAX_EXTRA) return -1;
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  for (ALevel &Lv : A->lev) {
    double *p = nullptr;
    if (hipMalloc(&p, Lv.pool * sizeof(double)) != hipSuccess) {
      std::fprintf(stderr, "pamr_amr: out of device memory for a new field\n");
      return -1;
    }
    PMG_HOST_CHECK(hipMemset(p, 0, Lv.pool * sizeof(double)));
    Lv.extra.push_back(p);
  }
  return PMG_F_EXTRA + A->nextra++;
}

long long pmg_amr_pool_points(pmg_amr *A, int L) {
  if (!A || L < 1 || L > A->nlev) return -1;
  return (long long)A->lev[L - 1].pool;
}

// the multi-field transfers below only enqueue on pmg_amr_stream; fields are
// grouped kMaxFields per launch (each field reads and writes only itself)
static int run_grouped(pmg_amr *A, const XPlan &P, int Ls, int Ld, int kind, int nf, const int *which) {
  for (int f0 = 0; f0 < nf; f0 += kMaxFields) {
    const int m = std::min(kMaxFields, nf - f0);
    int rc = run_plan(A, P, Ls, Ld, kind, which + f0, m);
    if (rc) return rc;
  }
  return 0;
}

int pmg_amr_sync_fields(pmg_amr *A, int L, int bdy, int nf, const int *which) {
  if (!A || L < 1 || L > A->nlev || nf < 0 || (nf > 0 && !which)) return -1;
  if (bdy != 0 && bdy != 1) return -2;
  for (int f = 0; f < nf; f++) if (!valid_field(A, which[f])) return -1;
  ALevel &Lv = A->lev[L - 1];
  return run_grouped(A, bdy ? Lv.sync1 : Lv.sync0, L, L, 0, nf, which);
}

int pmg_amr_sync_fields_coarsest(pmg_amr *A, int L, int nf, const int *which) {
  if (!A || L < 1 || L > A->nlev || nf < 0 || (nf > 0 && !which)) return -1;
  for (int f = 0; f < nf; f++) if (!valid_field(A, which[f])) return -1;
  ALevel &Lv = A->lev[L - 1];
  int nco = 0;
  for (const HGrid &g : Lv.g) nco += g.coarsest ? 1 : 0;
  if (nco == 0) return 0;
  return run_grouped(A, nco == (int)Lv.g.size() ? Lv.sync1 : Lv.sync1c, L, L, 0, nf, which);
}

int pmg_amr_inject_fields(pmg_amr *A, int Lf, int nf, const int *which, const int *ops) {
  if (!A || Lf < 2 || Lf > A->nlev || nf < 0 || (nf > 0 && (!which || !ops))) return -1;
  for (int f = 0; f < nf; f++) if (!valid_field(A, which[f]) || ops[f] < 1 || ops[f] > 3) return -1;
  for (int op = 1; op <= 3; op++) {
    std::vector<int> sel;
    for (int f = 0; f < nf; f++) if (ops[f] == op) sel.push_back(which[f]);
    int rc = run_grouped(A, A->lev[Lf - 2].inject, Lf, Lf - 1, op, (int)sel.size(), sel.data());
    if (rc) return rc;
  }
  return 0;
}

int pmg_amr_num_sweep_variants(void) { return kAmrVariants; }

int pmg_amr_set_level_sweep_config(pmg_amr *A, int L, int variant, int zchunk) {
  int tx, ty;
  if (!A || L < 1 || L > A->nlev || zchunk < 1 || !variant_txy(variant, tx, ty)) return -1;
  ALevel &Lv = A->lev[L - 1];
  PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
  const int ov = Lv.variant, oz = Lv.zchunk;
  Lv.variant = variant;
  Lv.zchunk = zchunk;
  int rc = build_tiles(A, Lv);
  if (!rc && std::max(Lv.nstiles, Lv.nptiles) + 64 > A->ws.cap) rc = -3;  // more tiles than norm slots
  if (rc) {
    Lv.variant = ov;
    Lv.zchunk = oz;
    if (build_tiles(A, Lv)) return -1;
  }
  // the fold table points at the tile ranges just rebuilt
  if (A->d_fold) {
    (void)hipFree(A->d_fold);
    A->d_fold = nullptr;
  }
  if (build_folds(A)) return -1;
  drop_graphs(A);
  return rc;
}

int pmg_amr_sweep_config(pmg_amr *A, int L, int *variant, int *zchunk) {
  if (!A || L < 1 || L > A->nlev) return -1;
  if (variant) *variant = A->lev[L - 1].variant;
  if (zchunk) *zchunk = A->lev[L - 1].zchunk;
  return 0;
}

int pmg_amr_fill_cmask(pmg_amr *A, int L, int which, int field) {
  if (!A || L < 1 || L > A->nlev || (which != 0 && which != 1) || !valid_field(A, field)) return -1;
  ALevel &Lv = A->lev[L - 1];
  double *dst = field_pool(Lv, field);
  if (!dst) return -1;
  if (!Lv.cm[which]) {
    std::vector<double> h(Lv.pool, 0.0);
    std::vector<uint8_t> m;
    for (int gi = 0; gi < (int)Lv.g.size(); gi++) {
      cmask_bytes(A, L, gi, which, m);
      for (size_t q = 0; q < m.size(); q++) h[Lv.g[gi].off + q] = m[q] ? 1.0 : 0.0;
    }
    if (hipMalloc(&Lv.cm[which], Lv.pool * sizeof(double)) != hipSuccess) {
      Lv.cm[which] = nullptr;
      return -1;
    }
    PMG_HOST_CHECK(hipMemcpy(Lv.cm[which], h.data(), Lv.pool * sizeof(double), hipMemcpyHostToDevice));
  }
  PMG_HOST_CHECK(hipMemcpyAsync(dst, Lv.cm[which], Lv.pool * sizeof(double), hipMemcpyDeviceToDevice, A->stream));
  return 0;
}

int pmg_amr_interp_fields(pmg_amr *A, int Lc, int nf, const int *which) {
  return pmg_amr_interp_fields_ord(A, Lc, 2, nf, which);
}

// PAMR_inject with excision: local_transfer's coarsen branch (src/transfer.c:
// 1126-1149) one transfer at a time in list order, each the device
// m_util_r8 routine over the transfer's index boxes with the fine grid's
// mask; fields inner, as the reference loops over gfns per transfer
int pmg_amr_inject_fields_ex(pmg_amr *A, int Lf, int nf, const int *which, const int *ops, int mask, double ex) {
  if (!A || Lf < 2 || Lf > A->nlev || nf < 0 || (nf > 0 && (!which || !ops)) || !valid_field(A, mask)) return -1;
  for (int f = 0; f < nf; f++) if (!valid_field(A, which[f]) || ops[f] < 1 || ops[f] > 3) return -1;
  ALevel &S = A->lev[Lf - 1], &D = A->lev[Lf - 2];
  const double *chr = field_pool(S, mask);
  for (const XJob &J : D.inject.raw) {
    const HGrid &gs = S.g[J.sg], &gd = D.g[J.dg];
    int i1[3], i2[3], cnt[3];
    for (int d = 0; d < 3; d++) {
      i2[d] = J.lo[d] + 1;
      cnt[d] = J.hi[d] - J.lo[d] + 1;
      i1[d] = J.sh[d] + 2 * J.lo[d] + 1;
    }
    for (int f = 0; f < nf; f++) {
      const double *src = field_pool(S, which[f]) + gs.off;
      double *dst = field_pool(D, which[f]) + gd.off;
      int rc;
      if (ops[f] == 1)
        rc = pmg_dmcopy3d(src, dst, gs.n[0], gs.n[1], gs.n[2], gd.n[0], gd.n[1], gd.n[2], i1[0], i1[1], i1[2], i2[0],
                          i2[1], i2[2], cnt[0], cnt[1], cnt[2], 2, 2, 2, 1, 1, 1, A->stream);
      else
        rc = (ops[f] == 2 ? pmg_dmhwr3d : pmg_dmfwr3d)(src, dst, gs.n[0], gs.n[1], gs.n[2], gd.n[0], gd.n[1], gd.n[2],
                                                      i1[0], i1[1], i1[2], i2[0], i2[1], i2[2], cnt[0], cnt[1],
                                                      cnt[2], 2, 2, 2, 1, 1, 1, chr + gs.off, ex, 1, A->stream);
      if (rc) return rc;
    }
  }
  return 0;
}

// PAMR_interp with excision: local_transfer's interpolation branch
// (src/transfer.c:1089-1122) one transfer at a time in list order: the
// buffered temporary interpolated with the source grid's mask, its part
// copied into the destination, then clear_ex of the whole source grid and
// the whole destination grid (later transfers from the same source read the
// cleared values, as in the reference)
int pmg_amr_interp_fields_ex(pmg_amr *A, int Lc, int ord, int nf, const int *which, int mask, double ex) {
  if (!A || Lc < 1 || Lc >= A->nlev || nf < 0 || (nf > 0 && !which) || !valid_field(A, mask)) return -1;
  if (ord != 2 && ord != 4) return -2;
  for (int f = 0; f < nf; f++) if (!valid_field(A, which[f])) return -1;
  ALevel &S = A->lev[Lc - 1], &D = A->lev[Lc];
  const double *chrs = field_pool(S, mask), *chrd = field_pool(D, mask);
  for (const XJob &J : D.interp.raw) {
    const HGrid &gs = S.g[J.sg], &gd = D.g[J.dg];
    int i1i[3], nii[3], nb[3], i2b[3], i2[3], cnt[3];
    size_t tsize = 1;
    for (int d = 0; d < 3; d++) {
      i1i[d] = J.t0[d] + 1;
      nii[d] = J.tn[d];
      nb[d] = (nii[d] - 1) * 2 + 1;
      i2[d] = J.lo[d] + 1;
      cnt[d] = J.hi[d] - J.lo[d] + 1;
      i2b[d] = J.sh[d] + J.lo[d] - 2 * J.t0[d] + 1;
      tsize *= (size_t)nb[d];
    }
    if (tsize > A->xtmp_cap) {
      PMG_HOST_CHECK(hipStreamSynchronize(A->stream));
      if (A->xtmp) PMG_HOST_CHECK(hipFree(A->xtmp));
      A->xtmp = nullptr;
      A->xtmp_cap = 0;
      PMG_HOST_CHECK(hipMalloc(&A->xtmp, tsize * sizeof(double)));
      A->xtmp_cap = tsize;
    }
    for (int f = 0; f < nf; f++) {
      double *src = field_pool(S, which[f]) + gs.off, *dst = field_pool(D, which[f]) + gd.off;
      int rc = pmg_dminterp3d(src, A->xtmp, gs.n[0], gs.n[1], gs.n[2], nb[0], nb[1], nb[2], i1i[0], i1i[1], i1i[2],
                              1, 1, 1, nii[0], nii[1], nii[2], 1, 1, 1, 2, 2, 2, ord, chrs + gs.off, ex, 1, A->stream);
      if (!rc)
        rc = pmg_dmcopy3d(A->xtmp, dst, nb[0], nb[1], nb[2], gd.n[0], gd.n[1], gd.n[2], i2b[0], i2b[1], i2b[2], i2[0],
                          i2[1], i2[2], cnt[0], cnt[1], cnt[2], 1, 1, 1, 1, 1, 1, A->stream);
      if (rc) return rc;
      const size_t ns = gs.size(), nd = gd.size();
      hipLaunchKernelGGL(k_clear_ex, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, A->stream, src,
                         chrs + gs.off, ex, ns);
      hipLaunchKernelGGL(k_clear_ex, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, A->stream, dst,
                         chrd + gd.off, ex, nd);
      PMG_HOST_CHECK(hipGetLastError());
    }
  }
  return 0;
}

int pmg_amr_interp_fields_ord(pmg_amr *A, int Lc, int ord, int nf, const int *which) {
  if (!A || Lc < 1 || Lc >= A->nlev || nf < 0 || (nf > 0 && !which)) return -1;
  if (ord != 2 && ord != 4) return -2;
  for (int f = 0; f < nf; f++) if (!valid_field(A, which[f])) return -1;
  return run_grouped(A, A->lev[Lc].interp, Lc, Lc + 1, ord == 4 ? 5 : 4, nf, which);
}

}  // extern "C"
