// pmg_band.hip -- the learned correction near excision boundaries.
//
// With excision, the reference's transfers (renormalised restriction,
// excised interpolation and clear_ex, src/m_util_r8.f:835-949, 368-421,
// src/transfer.c:930-935) leave the coarse-grid correction short in a thin
// layer of fine points next to each excised region: the error of a V-cycle
// there is what slows the whole iteration down.  The learned correction acts
// only in that layer (the "band": updatable points within Euclidean distance
// `depth` grid points of an excised point).  The band is covered by cubic
// cores of side C; a network application gathers one crop of side C + 6
// around every core that holds band points (3 = the receptive-field radius
// of three 3x3x3 convolutions, so the outputs inside a core do not see the
// crop's zero padding), runs the network on the crops as a batch of boxes,
// and adds its output to u at the band points of each core.
//
// Everything here works on plain device pointers; the solver (pmg_solver.hip)
// owns the band plan and drives it.  No reference counterpart.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pmg_common.h"
#include "pmg_ops.h"
#include "pamr_mg_gpu.h"

using namespace pmg;

namespace {

// crop side -> crop point (x fastest)
__device__ __forceinline__ void crop_xyz(size_t q, int side, int &x, int &y, int &z) {
  x = (int)(q % side);
  y = (int)((q / side) % side);
  z = (int)(q / ((size_t)side * side));
}

// out[c][q] = scale * src[p] for p = org[c] + q inside the grid, 0 outside;
// scale = mul / *norm when norm is given (a norm the device folded), else mul
template <typename TS>
__global__ __launch_bounds__(256) void k_band_gather(const TS *__restrict__ src, int nx, int ny, int nz,
                                                     const int *__restrict__ org, int side, float mul,
                                                     const double *__restrict__ norm, float *__restrict__ out) {
  const int c = blockIdx.y;
  const size_t n = (size_t)side * side * side;
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  int x, y, z;
  crop_xyz(q, side, x, y, z);
  const int i = org[3 * c] + x, j = org[3 * c + 1] + y, k = org[3 * c + 2] + z;
  float v = 0.f;
  if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) {
    const double s = norm ? (double)mul / *norm : (double)mul;
    v = (float)(s * (double)src[ix3(i, j, k, nx, ny)]);
  }
  out[(size_t)c * n + q] = v;
}

// u[p] += coef * y[c][q] at the band points of each crop's core (crop points
// at least `margin` from its faces); coef = mul * *norm when norm is given.
// u2, when given, receives the same value (the twin sweep buffer).
template <typename TU>
__global__ __launch_bounds__(256) void k_band_scatter(TU *__restrict__ u, TU *__restrict__ u2,
                                                      const float *__restrict__ y,
                                                      const uint8_t *__restrict__ band, int nx, int ny, int nz,
                                                      const int *__restrict__ org, int side, int margin, double mul,
                                                      const double *__restrict__ norm) {
  const int c = blockIdx.y;
  const int core = side - 2 * margin;
  const size_t n = (size_t)core * core * core;
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  int x, yy, z;
  crop_xyz(q, core, x, yy, z);
  const int i = org[3 * c] + margin + x, j = org[3 * c + 1] + margin + yy, k = org[3 * c + 2] + margin + z;
  if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) return;
  const size_t p = ix3(i, j, k, nx, ny);
  if (!band[p]) return;
  const double coef = norm ? mul * *norm : mul;
  const size_t qc = ((size_t)(z + margin) * side + (yy + margin)) * side + (x + margin);
  const TU v = u[p] + (TU)(coef * (double)y[(size_t)c * side * side * side + qc]);
  u[p] = v;
  if (u2) u2[p] = v;
}

// The FP64 residual hook r = L[u] (finest level, rhs = 0) at the updatable
// points of each core grown by one point, its negation in FP32 into neg32:
// after a band correction the defect changes only there.  Overlapping
// regions of neighbouring cores write identical values.
template <int OP, bool EX, bool P2>
__global__ __launch_bounds__(256) void k_band_resid(const double *__restrict__ u, const double *__restrict__ rho,
                                                    const double *__restrict__ a, const double *__restrict__ cc,
                                                    const uint8_t *__restrict__ chr, float *__restrict__ neg32,
                                                    int nx, int ny, int nz, const int *__restrict__ org, int side,
                                                    int margin, Spacing<double> sp) {
  const int c = blockIdx.y;
  const int w = side - 2 * margin + 2;
  const size_t n = (size_t)w * w * w;
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  int x, y, z;
  crop_xyz(q, w, x, y, z);
  const int i = org[3 * c] + margin - 1 + x, j = org[3 * c + 1] + margin - 1 + y, k = org[3 * c + 2] + margin - 1 + z;
  if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2 || k < 1 || k > nz - 2) return;
  const size_t p = ix3(i, j, k, nx, ny);
  if constexpr (EX) { if (chr[p]) return; }
  const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
  const double L = lop_at<double, OP, P2>(u, rho, a, cc, p, sy, sz, sp, (double *)nullptr);
  const double res = L - 0.0;
  neg32[p] = (float)(-res);
}

// training crops: x = xs * xsrc, t = ts * (ustar - u), w = band indicator
// (0 outside the grid)
__global__ __launch_bounds__(256) void k_band_pairs(const float *__restrict__ xsrc, float xs,
                                                    const double *__restrict__ ustar, const double *__restrict__ u,
                                                    double ts, const uint8_t *__restrict__ band, int nx, int ny,
                                                    int nz, const int *__restrict__ org, int side,
                                                    float *__restrict__ xo, float *__restrict__ to,
                                                    float *__restrict__ wo) {
  const int c = blockIdx.y;
  const size_t n = (size_t)side * side * side;
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  int x, y, z;
  crop_xyz(q, side, x, y, z);
  const int i = org[3 * c] + x, j = org[3 * c + 1] + y, k = org[3 * c + 2] + z;
  float xv = 0.f, tv = 0.f, wv = 0.f;
  if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) {
    const size_t p = ix3(i, j, k, nx, ny);
    xv = xs * xsrc[p];
    tv = (float)(ts * (ustar[p] - u[p]));
    wv = band[p] ? 1.f : 0.f;
  }
  const size_t o = (size_t)c * n + q;
  xo[o] = xv;
  to[o] = tv;
  wo[o] = wv;
}

// The FP32 residual hook r = L32[u] (source rho, finest rhs = 0) at crop
// point p, or 0 where p is outside the grid or not updatable.
template <int OP, bool EX, bool P2>
__device__ __forceinline__ float band_res32(const float *__restrict__ u, const float *__restrict__ rho,
                                            const float *__restrict__ a, const float *__restrict__ cc,
                                            const uint8_t *__restrict__ chr, int nx, int ny, int nz, int i, int j,
                                            int k, const Spacing<float> &sp) {
  if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2 || k < 1 || k > nz - 2) return 0.f;
  const size_t p = ix3(i, j, k, nx, ny);
  if constexpr (EX) { if (chr[p]) return 0.f; }
  const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
  const float L = lop_at<float, OP, P2>(u, rho, a, cc, p, sy, sz, sp, (float *)nullptr);
  return L - 0.0f;
}

// network input inside the FP32 V-cycle: out[c][q] = mul / *norm * r32(p)
template <int OP, bool EX, bool P2>
__global__ __launch_bounds__(256) void k_band_gather_res(const float *__restrict__ u, const float *__restrict__ rho,
                                                         const float *__restrict__ a, const float *__restrict__ cc,
                                                         const uint8_t *__restrict__ chr, int nx, int ny, int nz,
                                                         const int *__restrict__ org, int side, float mul,
                                                         const double *__restrict__ norm, Spacing<float> sp,
                                                         float *__restrict__ out) {
  const int c = blockIdx.y;
  const size_t n = (size_t)side * side * side;
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  int x, y, z;
  crop_xyz(q, side, x, y, z);
  const int i = org[3 * c] + x, j = org[3 * c + 1] + y, k = org[3 * c + 2] + z;
  const float r = band_res32<OP, EX, P2>(u, rho, a, cc, chr, nx, ny, nz, i, j, k, sp);
  out[(size_t)c * n + q] = (float)(((double)mul / *norm) * (double)r);
}

// training pairs inside the FP32 V-cycle: x = r32 / *norm, t = (ustar - u64 -
// u32) / *norm (the error of the cycle's current correction), w = band
template <int OP, bool EX, bool P2>
__global__ __launch_bounds__(256) void k_band_pairs_res(const float *__restrict__ u, const float *__restrict__ rho,
                                                        const float *__restrict__ a, const float *__restrict__ cc,
                                                        const uint8_t *__restrict__ chr,
                                                        const double *__restrict__ ustar,
                                                        const double *__restrict__ u64,
                                                        const uint8_t *__restrict__ band, int nx, int ny, int nz,
                                                        const int *__restrict__ org, int side,
                                                        const double *__restrict__ norm, Spacing<float> sp,
                                                        float *__restrict__ xo, float *__restrict__ to,
                                                        float *__restrict__ wo) {
  const int c = blockIdx.y;
  const size_t n = (size_t)side * side * side;
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  int x, y, z;
  crop_xyz(q, side, x, y, z);
  const int i = org[3 * c] + x, j = org[3 * c + 1] + y, k = org[3 * c + 2] + z;
  const double s = 1.0 / *norm;
  const float r = band_res32<OP, EX, P2>(u, rho, a, cc, chr, nx, ny, nz, i, j, k, sp);
  float tv = 0.f, wv = 0.f;
  if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) {
    const size_t p = ix3(i, j, k, nx, ny);
    tv = (float)(s * (ustar[p] - u64[p] - (double)u[p]));
    wv = band[p] ? 1.f : 0.f;
  }
  const size_t o = (size_t)c * n + q;
  xo[o] = (float)(s * (double)r);
  to[o] = tv;
  wo[o] = wv;
}

// FP32 spacing as the solver builds it for its FP32 cycles
Spacing<float> spacing32(double hx, double hy, double hz) {
  Spacing<float> sp;
  const float dx = (float)hx, dy = (float)hy, dz = (float)hz;
  sp.dx2 = dx * dx;
  sp.dy2 = dy * dy;
  sp.dz2 = dz * dz;
  sp.idx2 = 1.0f / sp.dx2;
  sp.idy2 = 1.0f / sp.dy2;
  sp.idz2 = 1.0f / sp.dz2;
  sp.jac = -2 / sp.dx2 - 2 / sp.dy2 - 2 / sp.dz2;
  sp.ijac = 1 / sp.jac;
  auto p2 = [](double v) { int e; return std::frexp(v, &e) == 0.5 && v > 1e-300 && v < 1e300; };
  sp.pow2 = p2((double)sp.dx2) && p2((double)sp.dy2) && p2((double)sp.dz2);
  return sp;
}

inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" {

// Band of updatable points next to the excised region (host, no GPU):
// band[p] = 1 when p is interior (1 <= i <= nx-2, likewise j, k), not
// excised (chr != ex) and within Euclidean distance `depth` (grid points)
// of an excised point; 0 elsewhere.  The nearest excised point of any
// non-excised point has a non-excised 6-neighbour (a step from it towards
// the point would otherwise be a nearer excised point), so only those
// surface points are expanded.  Returns the number of band points.
long long pmg_band_mask(const double *chr, int nx, int ny, int nz, double ex, int depth, unsigned char *band) {
  if (!chr || !band || nx < 3 || ny < 3 || nz < 3 || depth < 1 || depth > 16) return -1;
  const size_t n = (size_t)nx * ny * nz;
  std::fill(band, band + n, (unsigned char)0);
  const int d2 = depth * depth;
  long long count = 0;
  auto excised = [&](int i, int j, int k) { return chr[ix3(i, j, k, nx, ny)] == ex; };
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        if (!excised(i, j, k)) continue;
        bool surface = false;
        if (i > 0 && !excised(i - 1, j, k)) surface = true;
        if (i < nx - 1 && !excised(i + 1, j, k)) surface = true;
        if (j > 0 && !excised(i, j - 1, k)) surface = true;
        if (j < ny - 1 && !excised(i, j + 1, k)) surface = true;
        if (k > 0 && !excised(i, j, k - 1)) surface = true;
        if (k < nz - 1 && !excised(i, j, k + 1)) surface = true;
        if (!surface) continue;
        for (int dk = -depth; dk <= depth; dk++)
          for (int dj = -depth; dj <= depth; dj++)
            for (int di = -depth; di <= depth; di++) {
              if (di * di + dj * dj + dk * dk > d2) continue;
              const int a = i + di, b = j + dj, c = k + dk;
              if (a < 1 || a > nx - 2 || b < 1 || b > ny - 2 || c < 1 || c > nz - 2) continue;
              const size_t p = ix3(a, b, c, nx, ny);
              if (band[p] || chr[p] == ex) continue;
              band[p] = 1;
              count++;
            }
      }
  return count;
}

// Crop origins of the cores (side `core`, tiling the grid from index 0) that
// hold at least one band point: origin = core corner - margin.  org must hold
// 3 * (number of cores) ints; returns the number written.
int pmg_band_cores(const unsigned char *band, int nx, int ny, int nz, int core, int margin, int *org, int cap) {
  if (!band || !org || core < 1 || margin < 0) return -1;
  const int cx = (nx + core - 1) / core, cy = (ny + core - 1) / core, cz = (nz + core - 1) / core;
  int m = 0;
  for (int c = 0; c < cz; c++)
    for (int b = 0; b < cy; b++)
      for (int a = 0; a < cx; a++) {
        bool any = false;
        for (int k = c * core; k < std::min(nz, (c + 1) * core) && !any; k++)
          for (int j = b * core; j < std::min(ny, (b + 1) * core) && !any; j++)
            for (int i = a * core; i < std::min(nx, (a + 1) * core); i++)
              if (band[ix3(i, j, k, nx, ny)]) { any = true; break; }
        if (!any) continue;
        if (m >= cap) return -1;
        org[3 * m] = a * core - margin;
        org[3 * m + 1] = b * core - margin;
        org[3 * m + 2] = c * core - margin;
        m++;
      }
  return m;
}

// Device launchers (no synchronisation; org is a device array of 3*ncrop ints).
int pmg_band_gather_f32(const float *src, int nx, int ny, int nz, const int *org, int ncrop, int side, float mul,
                        const double *norm, float *out, void *stream) {
  if (!src || !org || !out || ncrop < 1 || side < 1) return -1;
  hipLaunchKernelGGL(k_band_gather<float>, dim3(nblk((size_t)side * side * side), ncrop), dim3(256), 0,
                     (hipStream_t)stream, src, nx, ny, nz, org, side, mul, norm, out);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int pmg_band_scatter(double *u, double *u2, float *u32, const float *y, const unsigned char *band, int nx, int ny,
                     int nz, const int *org, int ncrop, int side, int margin, double mul, const double *norm,
                     void *stream) {
  if (!y || !band || !org || ncrop < 1 || side <= 2 * margin || (!u && !u32)) return -1;
  const int core = side - 2 * margin;
  const dim3 gr(nblk((size_t)core * core * core), ncrop);
  if (u)
    hipLaunchKernelGGL(k_band_scatter<double>, gr, dim3(256), 0, (hipStream_t)stream, u, u2, y, band, nx, ny, nz,
                       org, side, margin, mul, norm);
  else
    hipLaunchKernelGGL(k_band_scatter<float>, gr, dim3(256), 0, (hipStream_t)stream, u32, (float *)nullptr, y, band,
                       nx, ny, nz, org, side, margin, mul, norm);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int pmg_band_resid(int op, int do_ex, const double *u, const double *rho, const double *a, const double *c,
                   const unsigned char *chr, float *neg32, int nx, int ny, int nz, const int *org, int ncrop, int side,
                   int margin, double dx, double dy, double dz, void *stream) {
  if (!u || !rho || !neg32 || !org || ncrop < 1 || (op == 1 && (!a || !c)) || (do_ex && !chr)) return -1;
  Spacing<double> sp;
  sp.dx2 = dx * dx;
  sp.dy2 = dy * dy;
  sp.dz2 = dz * dz;
  sp.idx2 = 1.0 / sp.dx2;
  sp.idy2 = 1.0 / sp.dy2;
  sp.idz2 = 1.0 / sp.dz2;
  sp.jac = -2 / sp.dx2 - 2 / sp.dy2 - 2 / sp.dz2;
  sp.ijac = 1 / sp.jac;
  auto p2of = [](double v) { int e; return std::frexp(v, &e) == 0.5; };
  sp.pow2 = p2of(sp.dx2) && p2of(sp.dy2) && p2of(sp.dz2);
  const int w = side - 2 * margin + 2;
  const dim3 gr(nblk((size_t)w * w * w), ncrop);
  hipStream_t st = (hipStream_t)stream;
#define PMG_BR(OPV, EXV, P2V)                                                                                    \
  hipLaunchKernelGGL((k_band_resid<OPV, EXV, P2V>), gr, dim3(256), 0, st, u, rho, a, c, (const uint8_t *)chr, neg32, \
                     nx, ny, nz, org, side, margin, sp)
  if (op == 0) {
    if (do_ex) { if (sp.pow2) PMG_BR(0, true, true); else PMG_BR(0, true, false); }
    else { if (sp.pow2) PMG_BR(0, false, true); else PMG_BR(0, false, false); }
  } else {
    if (do_ex) { if (sp.pow2) PMG_BR(1, true, true); else PMG_BR(1, true, false); }
    else { if (sp.pow2) PMG_BR(1, false, true); else PMG_BR(1, false, false); }
  }
#undef PMG_BR
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int pmg_band_gather_res32(int op, int do_ex, const float *u, const float *rho, const float *a, const float *c,
                          const unsigned char *chr, int nx, int ny, int nz, const int *org, int ncrop, int side,
                          float mul, const double *norm, double dx, double dy, double dz, float *out, void *stream) {
  if (!u || !rho || !org || !norm || !out || ncrop < 1 || (op == 1 && (!a || !c)) || (do_ex && !chr)) return -1;
  const Spacing<float> sp = spacing32(dx, dy, dz);
  const dim3 gr(nblk((size_t)side * side * side), ncrop);
  hipStream_t st = (hipStream_t)stream;
  const uint8_t *ch = (const uint8_t *)chr;
#define PMG_BG(OPV, EXV, P2V)                                                                                     \
  hipLaunchKernelGGL((k_band_gather_res<OPV, EXV, P2V>), gr, dim3(256), 0, st, u, rho, a, c, ch, nx, ny, nz, org, side, \
                     mul, norm, sp, out)
  if (op == 0) {
    if (do_ex) { if (sp.pow2) PMG_BG(0, true, true); else PMG_BG(0, true, false); }
    else { if (sp.pow2) PMG_BG(0, false, true); else PMG_BG(0, false, false); }
  } else {
    if (do_ex) { if (sp.pow2) PMG_BG(1, true, true); else PMG_BG(1, true, false); }
    else { if (sp.pow2) PMG_BG(1, false, true); else PMG_BG(1, false, false); }
  }
#undef PMG_BG
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int pmg_band_pairs_res32(int op, int do_ex, const float *u, const float *rho, const float *a, const float *c,
                         const unsigned char *chr, const double *ustar, const double *u64, const unsigned char *band,
                         int nx, int ny, int nz, const int *org, int ncrop, int side, const double *norm, double dx,
                         double dy, double dz, float *xo, float *to, float *wo, void *stream) {
  if (!u || !rho || !ustar || !u64 || !band || !org || !norm || !xo || !to || !wo || ncrop < 1 ||
      (op == 1 && (!a || !c)) || (do_ex && !chr))
    return -1;
  const Spacing<float> sp = spacing32(dx, dy, dz);
  const dim3 gr(nblk((size_t)side * side * side), ncrop);
  hipStream_t st = (hipStream_t)stream;
  const uint8_t *ch = (const uint8_t *)chr, *bd = (const uint8_t *)band;
#define PMG_BP(OPV, EXV, P2V)                                                                                    \
  hipLaunchKernelGGL((k_band_pairs_res<OPV, EXV, P2V>), gr, dim3(256), 0, st, u, rho, a, c, ch, ustar, u64, bd, nx, ny, \
                     nz, org, side, norm, sp, xo, to, wo)
  if (op == 0) {
    if (do_ex) { if (sp.pow2) PMG_BP(0, true, true); else PMG_BP(0, true, false); }
    else { if (sp.pow2) PMG_BP(0, false, true); else PMG_BP(0, false, false); }
  } else {
    if (do_ex) { if (sp.pow2) PMG_BP(1, true, true); else PMG_BP(1, true, false); }
    else { if (sp.pow2) PMG_BP(1, false, true); else PMG_BP(1, false, false); }
  }
#undef PMG_BP
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int pmg_band_pairs(const float *xsrc, float xs, const double *ustar, const double *u, double ts,
                   const unsigned char *band, int nx, int ny, int nz, const int *org, int ncrop, int side, float *xo,
                   float *to, float *wo, void *stream) {
  if (!xsrc || !ustar || !u || !band || !org || !xo || !to || !wo || ncrop < 1 || side < 1) return -1;
  hipLaunchKernelGGL(k_band_pairs, dim3(nblk((size_t)side * side * side), ncrop), dim3(256), 0,
                     (hipStream_t)stream, xsrc, xs, ustar, u, ts, (const uint8_t *)band, nx, ny, nz, org, side, xo,
                     to, wo);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
