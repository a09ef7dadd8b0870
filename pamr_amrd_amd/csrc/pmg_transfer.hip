// pmg_transfer.hip -- device versions of the m_util_r8 transfer routines.
//
// Each extern "C" entry point takes exactly the arguments of the Fortran
// routine it replaces (src/m_util_r8.f; prototypes include/m_util_r8.h),
// by value, with u1/u2/chr1 being device pointers, plus a HIP stream.  The
// silent early returns of the Fortran bounds checks are kept (return 0,
// nothing written); launch errors return -1.
#include <hip/hip_runtime.h>

#include "pamr_mg_gpu.h"

#include "pmg_ops.h"

using namespace pmg;

namespace {

constexpr int NT = 256;

inline unsigned nblk(size_t n) { return (unsigned)((n + NT - 1) / NT); }

// dmcopy3d, src/m_util_r8.f:20-50
template <typename T>
__global__ __launch_bounds__(NT) void k_copy3d(const T *u1, T *u2, int nx1, int ny1, int nx2,
                                               int ny2, int i1, int j1, int k1, int i2, int j2,
                                               int k2, int ni, int nj, int nk, int si1, int sj1,
                                               int sk1, int si2, int sj2, int sk2) {
  size_t q = (size_t)blockIdx.x * NT + threadIdx.x;
  if (q >= (size_t)ni * nj * nk) return;
  int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
  u2[ix3(i2 - 1 + i * si2, j2 - 1 + j * sj2, k2 - 1 + k * sk2, nx2, ny2)] =
      u1[ix3(i1 - 1 + i * si1, j1 - 1 + j * sj1, k1 - 1 + k * sk1, nx1, ny1)];
}

// dmfwr3d / dmhwr3d over an ni x nj x nk target segment
template <typename T, int KIND, bool EXC>
__global__ __launch_bounds__(NT) void k_restrict3d(const T *u1, T *u2, const T *chr1, T ex,
                                                   int nx1, int ny1, int nz1, int nx2, int ny2,
                                                   int i1, int j1, int k1, int i2, int j2,
                                                   int k2, int ni, int nj, int nk, int si1,
                                                   int sj1, int sk1, int si2, int sj2,
                                                   int sk2) {
  size_t q = (size_t)blockIdx.x * NT + threadIdx.x;
  if (q >= (size_t)ni * nj * nk) return;
  int i = (int)(q % ni), j = (int)((q / ni) % nj), k = (int)(q / ((size_t)ni * nj));
  int i0 = i1 + i * si1, j0 = j1 + j * sj1, k0 = k1 + k * sk1;
  T v;
  if constexpr (KIND == 3) {
    if constexpr (EXC) v = fw_point_ex<T>(u1, chr1, ex, nx1, ny1, nz1, i0, j0, k0, i + 1);
    else v = fw_point<T>(u1, nx1, ny1, nz1, i0, j0, k0);
  } else {
    v = hw_point<T, EXC>(u1, chr1, ex, nx1, ny1, nz1, i0, j0, k0);
  }
  u2[ix3(i2 - 1 + i * si2, j2 - 1 + j * sj2, k2 - 1 + k * sk2, nx2, ny2)] = v;
}

// dminterp3d with ord = 2 and no excision: the three line phases of
// src/m_util_r8.f:518-634 composed per target point.  Along each axis the
// phase value is c0*v(a) + c1*v(a+1) (dvi2q1d, :350-357) with
// sg=(r*1.0)/rho, c0=1-sg, c1=sg, applied x first, then y, then z, on the
// previous phase's values -- the same IEEE operations in the same order.
template <typename T>
__device__ __forceinline__ T interp_coef(int r, int rho, T *c1) {
  double sg = (r * 1.0) / rho;
  *c1 = (T)sg;
  return (T)(1 - sg);
}

template <typename T>
__global__ __launch_bounds__(NT) void k_interp3d(const T *u1, T *u2, int nx1, int ny1, int nx2,
                                                 int ny2, int i1, int j1, int k1, int i2,
                                                 int j2, int k2, int ni, int nj, int nk,
                                                 int si1, int sj1, int sk1, int irho, int jrho,
                                                 int krho) {
  const int mi = (ni - 1) * irho + 1, mj = (nj - 1) * jrho + 1, mk = (nk - 1) * krho + 1;
  size_t q = (size_t)blockIdx.x * NT + threadIdx.x;
  if (q >= (size_t)mi * mj * mk) return;
  int fi = (int)(q % mi), fj = (int)((q / mi) % mj), fk = (int)(q / ((size_t)mi * mj));
  int ai = fi / irho, ri = fi % irho, aj = fj / jrho, rj = fj % jrho, ak = fk / krho,
      rk = fk % krho;
  auto U = [&](int a, int b, int c) -> T {
    return u1[ix3(i1 - 1 + a * si1, j1 - 1 + b * sj1, k1 - 1 + c * sk1, nx1, ny1)];
  };
  T cx1, cy1, cz1;
  T cx0 = interp_coef<T>(ri, irho, &cx1), cy0 = interp_coef<T>(rj, jrho, &cy1),
    cz0 = interp_coef<T>(rk, krho, &cz1);
  auto X = [&](int b, int c) -> T {
    return ri == 0 ? U(ai, b, c) : cx0 * U(ai, b, c) + cx1 * U(ai + 1, b, c);
  };
  auto Y = [&](int c) -> T { return rj == 0 ? X(aj, c) : cy0 * X(aj, c) + cy1 * X(aj + 1, c); };
  T v = rk == 0 ? Y(ak) : cz0 * Y(ak) + cz1 * Y(ak + 1);
  u2[ix3(i2 - 1 + fi, j2 - 1 + fj, k2 - 1 + fk, nx2, ny2)] = v;
}

// ---- fourth-order interpolation (dminterp3d, ord = 4, no excision) ------
// (vi4_point, one point of dvi4q1d, is in pmg_ops.h)

// The three line phases of dminterp3d (src/m_util_r8.f:518-634), one launch
// each, one thread per written point.  A phase writes only points off the
// anchor lattice of its axis and reads only points on it (inside or just
// outside the segment), so its points are independent; phases run in the
// reference's order x, y, z on u2 after the common points were copied.
template <int AX>
__global__ __launch_bounds__(NT) void k_interp4(double *u2, int nx2, int ny2, int nz2, int i2, int j2, int k2, int ni,
                                                int nj, int nk, int irho, int jrho, int krho) {
  const int mi = (ni - 1) * irho + 1, mj = (nj - 1) * jrho + 1, mk = (nk - 1) * krho + 1;
  const size_t q = (size_t)blockIdx.x * NT + threadIdx.x;
  int fi, fj, fk;  // offsets from (i2, j2, k2); the phase's own axis is the fine one
  if (AX == 0) {
    if (q >= (size_t)mi * nj * nk) return;
    fi = (int)(q % mi);
    const int line = (int)(q / mi);
    fj = (line % nj) * jrho;
    fk = (line / nj) * krho;
    if (fi % irho == 0) return;
    auto A = [&](int a) -> double { return u2[ix3(i2 - 1 + a * irho, j2 - 1 + fj, k2 - 1 + fk, nx2, ny2)]; };
    u2[ix3(i2 - 1 + fi, j2 - 1 + fj, k2 - 1 + fk, nx2, ny2)] = vi4_point(ni, irho, i2, nx2, fi, A);
  } else if (AX == 1) {
    if (q >= (size_t)mi * mj * nk) return;
    fj = (int)(q % mj);
    const size_t line = q / mj;
    fi = (int)(line % mi);
    fk = (int)(line / mi) * krho;
    if (fj % jrho == 0) return;
    auto A = [&](int a) -> double { return u2[ix3(i2 - 1 + fi, j2 - 1 + a * jrho, k2 - 1 + fk, nx2, ny2)]; };
    u2[ix3(i2 - 1 + fi, j2 - 1 + fj, k2 - 1 + fk, nx2, ny2)] = vi4_point(nj, jrho, j2, ny2, fj, A);
  } else {
    if (q >= (size_t)mi * mj * mk) return;
    fk = (int)(q / ((size_t)mi * mj));
    const size_t rest = q % ((size_t)mi * mj);
    fi = (int)(rest % mi);
    fj = (int)(rest / mi);
    if (fk % krho == 0) return;
    auto A = [&](int a) -> double { return u2[ix3(i2 - 1 + fi, j2 - 1 + fj, k2 - 1 + a * krho, nx2, ny2)]; };
    u2[ix3(i2 - 1 + fi, j2 - 1 + fj, k2 - 1 + fk, nx2, ny2)] = vi4_point(nk, krho, k2, nz2, fk, A);
  }
}

// ---- excision interpolation: one thread per line (dvi2q1d_ex) ----------
// Line view: v(q) for q = 1..n1 (1-based) at stride sv elements.
template <typename T, typename CHR>
__device__ void line_vi2q1d_ex(T *v, int i1, int n1, int n, int irho, int s1, int s2,
                               const CHR &chr, int i3, int s3, T ex) {
#define LV(q) v[((size_t)(q)-1)]
#define LC(q) chr((q)-1)
  for (int i = 1; i <= n; i++) {
    int i10 = (i - 1) * s1 + i1, i30 = (i - 1) * s3 + i3;
    if (LC(i30) == ex) {
      if (i10 > 3 * s1 && LC(i30 - s3) != ex && LC(i30 - 2 * s3) != ex && LC(i30 - 3 * s3) != ex)
        LV(i10) = 3 * (LV(i10 - s1) - LV(i10 - 2 * s1)) + LV(i10 - 3 * s1);
      else if (i10 <= (n1 - 3 * s1) && LC(i30 + s3) != ex && LC(i30 + 2 * s3) != ex &&
               LC(i30 + 3 * s3) != ex)
        LV(i10) = 3 * (LV(i10 + s1) - LV(i10 + 2 * s1)) + LV(i10 + 3 * s1);
      else if (i10 > 2 * s1 && LC(i30 - s3) != ex && LC(i30 - 2 * s3) != ex)
        LV(i10) = 2 * LV(i10 - s1) - LV(i10 - 2 * s1);
      else if (i10 <= (n1 - 2 * s1) && LC(i30 + s3) != ex && LC(i30 + 2 * s3) != ex)
        LV(i10) = 2 * LV(i10 + s1) - LV(i10 + 2 * s1);
      else if (i10 > s1 && LC(i30 - s3) != ex)
        LV(i10) = LV(i10 - s1);
      else if (i10 <= (n1 - s1) && LC(i30 + s3) != ex)
        LV(i10) = LV(i10 + s1);
    }
    // v2(irho*(i-1)*s2+i2) = v1(i10): in place with s1 == irho*s2, i1 == i2
  }
  for (int j = 1; j <= irho - 1; j++) {
    double sg = (j * 1.0) / irho;
    T c0 = (T)(1 - sg), c1 = (T)sg;
    for (int i = 0; i <= n - 2; i++) LV((irho * i + j) * s2 + i1) = c0 * LV(i * s1 + i1) + c1 * LV((i + 1) * s1 + i1);
  }
#undef LV
#undef LC
}

// dvi4q1d_ex (src/m_util_r8.f:146-319) on one line, in place (v1 = v2 with
// s1 == irho*s2, so the common-point copy is a no-op): the extension of the
// anchor range by up to 3 anchors each side, the dvi2q1d_ex pass when a
// non-excised run is shorter than 4 anchors, extrapolation into excised
// anchors, then the first / last / interior cubic cells -- in the
// reference's order, with its conditions as written.
template <typename T, typename CHR>
__device__ void line_vi4q1d_ex(T *v, int i1, int n1, int n, int irho, int s1, int s2, const CHR &chr, int i3,
                               int s3, T ex) {
#define LV(q) v[((size_t)(q)-1)]
#define LC(q) chr((q)-1)
  const double half = 0.5, mhalf = -0.5, sixth = 0.1666666666666667, msixth = -0.1666666666666667;
  int i = 3, i1s = i1;
  while (i > 0 && i1s > s1) {
    i1s -= s1;
    i--;
  }
  i = 3;
  int i1e = i1 + (n - 1) * s1;
  while (i > 0 && i1e < (n1 - s1)) {
    i1e += s1;
    i--;
  }
  int cn = 0;
  bool init_2 = false;
  for (i = i1s; i <= i1e; i += s1) {
    if (LC((i - i1) * s3 / s1 + i3) == ex || i == i1e) {
      if (cn > 0 && cn < 4) init_2 = true;
      else cn = 0;
    } else {
      cn++;
    }
  }
  if (init_2) line_vi2q1d_ex<T>(v, i1, n1, n, irho, s1, s2, chr, i3, s3, ex);
  for (i = i1s; i <= i1e; i += s1) {
    if (LC(i3 + s3 * (i - i1) / s1) == ex) {
      if (i >= (i1s + 3 * s1) && LC(i3 + s3 * (i - s1 - i1) / s1) != ex && LC(i3 + s3 * (i - 2 * s1 - i1) / s1) != ex &&
          LC(i3 + s3 * (i - 3 * s1 - i1) / s1) != ex) {
        if (i >= (i1s + 4 * s1) && LC(i3 + s3 * (i - 4 * s1 - i1) / s1) != ex)
          LV(i) = 4 * (LV(i - s1) + LV(i - 3 * s1)) - 6 * LV(i - 2 * s1) - LV(i - 4 * s1);
        else
          LV(i) = 3 * (LV(i - s1) - LV(i - 2 * s1)) + LV(i - 3 * s1);
      } else if (i <= (i1e - 3 * s1) && LC(i3 + s3 * (i + s1 - i1) / s1) != ex &&
                 LC(i3 + s3 * (i + 2 * s1 - i1) / s1) != ex && LC(i3 + s3 * (i + 3 * s1 - i1) / s1) != ex) {
        if (i <= (i1s - 4 * s1) && LC(i3 + s3 * (i + 4 * s1 - i1) / s1) != ex)
          LV(i) = 4 * (LV(i + s1) + LV(i + 3 * s1)) - 6 * LV(i + 2 * s1) - LV(i + 4 * s1);
        else
          LV(i) = 3 * (LV(i + s1) - LV(i + 2 * s1)) + LV(i + 3 * s1);
      }
    }
  }
  auto coef = [&](double sg, double *c) {
    c[0] = msixth * (sg - 1.0) * (sg - 2.0) * (sg - 3.0);
    c[1] = half * sg * (sg - 2.0) * (sg - 3.0);
    c[2] = mhalf * sg * (sg - 1.0) * (sg - 3.0);
    c[3] = sixth * sg * (sg - 1.0) * (sg - 2.0);
  };
  double c[4];
  for (i = 1; i <= n - 1; i++) {  // first cells of non-excised runs
    const int i10 = (i - 1) * s1 + i1, i30 = (i - 1) * s3 + i3, j0 = (i - 1) * irho * s2 + i1;
    if (i10 <= (i1e - 3 * s1) && LC(i30 + s3) != ex && LC(i30 + 2 * s3) != ex &&
        ((i == 1 && i10 == i1s) || (i10 >= (i1s + s1) && LC(i30) != ex && LC(i30 - s3) == ex) || LC(i30) == ex)) {
      for (int j = 1; j <= irho - 1; j++) {
        coef((j * 1.0) / irho, c);
        LV(j * s2 + j0) = c[0] * LV(i10) + c[1] * LV(s1 + i10) + c[2] * LV(2 * s1 + i10) + c[3] * LV(3 * s1 + i10);
      }
    }
  }
  for (i = 1; i <= n - 1; i++) {  // last cells of non-excised runs
    const int i10 = (i - 1) * s1 + i1, i30 = (i - 1) * s3 + i3, j0 = (i - 1) * irho * s2 + i1;
    if (i10 >= (i1s + 2 * s1) && LC(i30 - s3) != ex &&
        ((i == (n - 1) && i10 == (i1e - s1)) || (i10 <= (i1e - s1) && LC(i30) != ex && LC(i30 + s3) == ex))) {
      for (int j = 1; j <= irho - 1; j++) {
        coef(2.0 + (j * 1.0) / irho, c);
        LV(j * s2 + j0) = c[0] * LV(i10 - 2 * s1) + c[1] * LV(i10 - s1) + c[2] * LV(i10) + c[3] * LV(i10 + s1);
      }
    }
  }
  for (i = 1; i <= n; i++) {  // interior cells
    const int i10 = (i - 1) * s1 + i1, i30 = (i - 1) * s3 + i3, j0 = (i - 1) * irho * s2 + i1;
    if (LC(i30) != ex && i10 >= (i1s + s1) && LC(i30 - s3) != ex && i10 <= (i1e - 2 * s1) && LC(i30 + s3) != ex) {
      for (int j = 1; j <= irho - 1; j++) {
        coef(1.0 + (j * 1.0) / irho, c);
        LV(j0 + j * s2) = c[0] * LV(i10 - s1) + c[1] * LV(i10) + c[2] * LV(i10 + s1) + c[3] * LV(i10 + 2 * s1);
      }
    }
  }
#undef LV
#undef LC
}

// chrt of src/m_util_r8.f:426-456 evaluated on demand: init_chrt copies
// column c0, each and_chrt folds one more column (ex only if both are ex).
template <typename T>
struct ChrtCols {
  const T *c[4];
  size_t stride;
  T ex;
  __device__ T operator()(int m) const {
    size_t off = (size_t)m * stride;
    T v = c[0][off];
    for (int q = 1; q < 4; q++)
      if (c[q]) v = (v == ex && c[q][off] == ex) ? ex : ex + 1;
    return v;
  }
};

// x lines through anchor (j,k): in place on u2
template <typename T, int ORD>
__global__ void k_interp_ex_x(T *u2, const T *chr1, T ex, int nx1, int ny1, int nx2, int ny2,
                              int i1, int j1, int k1, int i2, int j2, int k2, int ni, int nj,
                              int nk, int si1, int sj1, int sk1, int irho, int jrho, int krho) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nj * nk) return;
  int jj = q % nj, kk = q / nj;
  int j = j2 + jj * jrho, k = k2 + kk * krho;
  T *line = u2 + ix3(0, j - 1, k - 1, nx2, ny2);
  ChrtCols<T> cl{{chr1 + ix3(0, j1 + sj1 * (j - j2) / jrho - 1, k1 + sk1 * (k - k2) / krho - 1,
                             nx1, ny1), nullptr, nullptr, nullptr}, 1, ex};
  if (ORD == 4) line_vi4q1d_ex<T>(line, i2, nx2, ni, irho, irho, 1, cl, i1, si1, ex);
  else line_vi2q1d_ex<T>(line, i2, nx2, ni, irho, irho, 1, cl, i1, si1, ex);
}

// y lines for every fine i at anchor k; chrt staged in a per-thread column
template <typename T, int ORD>
__global__ void k_interp_ex_y(T *u2, const T *chr1, T ex, int nx1, int ny1,
                              int nx2, int ny2, int i1, int j1, int k1, int i2, int j2, int k2,
                              int ni, int nj, int nk, int si1, int sj1, int sk1, int irho,
                              int jrho, int krho) {
  const int mi = (ni - 1) * irho + 1;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mi * nk) return;
  int ii = q % mi, kk = q / mi;
  int i = i2 + ii, k = k2 + kk * krho;
  int ic = i1 + si1 * ((i - i2) / irho), kc = k1 + sk1 * ((k - k2) / krho);
  ChrtCols<T> ch{{chr1 + ix3(ic - 1, 0, kc - 1, nx1, ny1),
                  ((i - i2) % irho != 0) ? chr1 + ix3(ic - 1 + si1, 0, kc - 1, nx1, ny1) : nullptr,
                  nullptr, nullptr},
                 (size_t)nx1, ex};
  T *col = u2 + ix3(i - 1, 0, k - 1, nx2, ny2);
  if (ORD == 4)
    line_vi4q1d_ex<T>(col, 1 + (j2 - 1) * nx2, nx2 * (ny2 - 1) + 1, nj, jrho, jrho * nx2, nx2, ch, j1, sj1, ex);
  else
    line_vi2q1d_ex<T>(col, 1 + (j2 - 1) * nx2, nx2 * (ny2 - 1) + 1, nj, jrho, jrho * nx2, nx2, ch,
                      j1, sj1, ex);
}

// z lines for every fine (i,j)
template <typename T, int ORD>
__global__ void k_interp_ex_z(T *u2, const T *chr1, T ex, int nx1, int ny1,
                              int nz1, int nx2, int ny2, int nz2, int i1, int j1, int k1, int i2,
                              int j2, int k2, int ni, int nj, int nk, int si1, int sj1, int sk1,
                              int irho, int jrho, int krho) {
  const int mi = (ni - 1) * irho + 1, mj = (nj - 1) * jrho + 1;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mi * mj) return;
  int ii = q % mi, jj = q / mi;
  int i = i2 + ii, j = j2 + jj;
  int ic = i1 + si1 * ((i - i2) / irho), jc = j1 + sj1 * ((j - j2) / jrho);
  bool oi = (i - i2) % irho != 0, oj = (j - j2) % jrho != 0;
  ChrtCols<T> ch{{chr1 + ix3(ic - 1, jc - 1, 0, nx1, ny1),
                  oi ? chr1 + ix3(ic - 1 + si1, jc - 1, 0, nx1, ny1) : nullptr,
                  oj ? chr1 + ix3(ic - 1, jc - 1 + sj1, 0, nx1, ny1) : nullptr,
                  (oi && oj) ? chr1 + ix3(ic - 1 + si1, jc - 1 + sj1, 0, nx1, ny1) : nullptr},
                 (size_t)nx1 * ny1, ex};
  (void)nz1;
  T *col = u2 + ix3(i - 1, j - 1, 0, nx2, ny2);
  const int s2 = nx2 * ny2;
  if (ORD == 4)
    line_vi4q1d_ex<T>(col, (k2 - 1) * s2 + 1, s2 * (nz2 - 1) + 1, nk, krho, krho * s2, s2, ch, k1, sk1, ex);
  else
    line_vi2q1d_ex<T>(col, (k2 - 1) * s2 + 1, s2 * (nz2 - 1) + 1, nk, krho, krho * s2, s2, ch, k1,
                      sk1, ex);
}

}  // namespace

extern "C" {

int pmg_dmcopy3d(const double *u1, double *u2, int nx1, int ny1, int nz1, int nx2, int ny2,
                 int nz2, int i1, int j1, int k1, int i2, int j2, int k2, int ni, int nj,
                 int nk, int si1, int sj1, int sk1, int si2, int sj2, int sk2,
                 hipStream_t stream) {
  if (i1 < 1 || (i1 + (ni - 1) * si1) > nx1) return 0;
  if (i2 < 1 || (i2 + (ni - 1) * si2) > nx2) return 0;
  if (j1 < 1 || (j1 + (nj - 1) * sj1) > ny1) return 0;
  if (j2 < 1 || (j2 + (nj - 1) * sj2) > ny2) return 0;
  if (k1 < 1 || (k1 + (nk - 1) * sk1) > nz1) return 0;
  if (k2 < 1 || (k2 + (nk - 1) * sk2) > nz2) return 0;
  size_t n = (size_t)ni * nj * nk;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_copy3d<double>, dim3(nblk(n)), dim3(NT), 0, stream, u1, u2, nx1, ny1,
                     nx2, ny2, i1, j1, k1, i2, j2, k2, ni, nj, nk, si1, sj1, sk1, si2, sj2, sk2);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

static int restrict_launch(int kind, const double *u1, double *u2, int nx1, int ny1, int nz1,
                           int nx2, int ny2, int nz2, int i1, int j1, int k1, int i2, int j2,
                           int k2, int ni, int nj, int nk, int si1, int sj1, int sk1, int si2,
                           int sj2, int sk2, const double *chr1, double ex, int do_ex,
                           hipStream_t stream) {
  if (i1 < 1 || (i1 + (ni - 1) * si1) > nx1) return 0;
  if (i2 < 1 || (i2 + (ni - 1) * si2) > nx2) return 0;
  if (j1 < 1 || (j1 + (nj - 1) * sj1) > ny1) return 0;
  if (j2 < 1 || (j2 + (nj - 1) * sj2) > ny2) return 0;
  if (k1 < 1 || (k1 + (nk - 1) * sk1) > nz1) return 0;
  if (k2 < 1 || (k2 + (nk - 1) * sk2) > nz2) return 0;
  size_t n = (size_t)ni * nj * nk;
  if (n == 0) return 0;
  const bool ex1 = (do_ex == 1);
  if (ex1 && !chr1) return -1;
#define PMG_RL(K, E)                                                                           \
  hipLaunchKernelGGL((k_restrict3d<double, K, E>), dim3(nblk(n)), dim3(NT), 0, stream, u1, u2, \
                     chr1, ex, nx1, ny1, nz1, nx2, ny2, i1, j1, k1, i2, j2, k2, ni, nj, nk,   \
                     si1, sj1, sk1, si2, sj2, sk2)
  if (kind == 3) {
    if (ex1) PMG_RL(3, true); else PMG_RL(3, false);
  } else {
    if (ex1) PMG_RL(2, true); else PMG_RL(2, false);
  }
#undef PMG_RL
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

int pmg_dmfwr3d(const double *u1, double *u2, int nx1, int ny1, int nz1, int nx2, int ny2,
                int nz2, int i1, int j1, int k1, int i2, int j2, int k2, int ni, int nj, int nk,
                int si1, int sj1, int sk1, int si2, int sj2, int sk2, const double *chr1,
                double ex, int do_ex, hipStream_t stream) {
  return restrict_launch(3, u1, u2, nx1, ny1, nz1, nx2, ny2, nz2, i1, j1, k1, i2, j2, k2, ni,
                         nj, nk, si1, sj1, sk1, si2, sj2, sk2, chr1, ex, do_ex, stream);
}

int pmg_dmhwr3d(const double *u1, double *u2, int nx1, int ny1, int nz1, int nx2, int ny2,
                int nz2, int i1, int j1, int k1, int i2, int j2, int k2, int ni, int nj, int nk,
                int si1, int sj1, int sk1, int si2, int sj2, int sk2, const double *chr1,
                double ex, int do_ex, hipStream_t stream) {
  return restrict_launch(2, u1, u2, nx1, ny1, nz1, nx2, ny2, nz2, i1, j1, k1, i2, j2, k2, ni,
                         nj, nk, si1, sj1, sk1, si2, sj2, sk2, chr1, ex, do_ex, stream);
}

int pmg_dminterp3d(double *u1, double *u2, int nx1, int ny1, int nz1, int nx2, int ny2, int nz2,
                   int i1, int j1, int k1, int i2, int j2, int k2, int ni, int nj, int nk,
                   int si1, int sj1, int sk1, int irho, int jrho, int krho, int ord,
                   const double *chr1, double ex, int do_ex, hipStream_t stream) {
  (void)u1;
  if (i1 < 1 || (i1 + ni - 1) * si1 > nx1) return 0;
  if (i2 < 1 || (i2 + (ni - 1) * irho) > nx2) return 0;
  if (j1 < 1 || (j1 + nj - 1) * sj1 > ny1) return 0;
  if (j2 < 1 || (j2 + (nj - 1) * jrho) > ny2) return 0;
  if (k1 < 1 || (k1 + nk - 1) * sk1 > nz1) return 0;
  if (k2 < 1 || (k2 + (nk - 1) * krho) > nz2) return 0;
  if (ord != 2 && ord != 4) {
    std::fprintf(stderr, "pmg_dminterp3d: only ord=2,4 are supported (got %d)\n", ord);
    return -2;
  }
  if (ord == 4 && do_ex == 1 && (irho > 20 || jrho > 20 || krho > 20)) {
    std::fprintf(stderr, "pmg_dminterp3d: dvi4q1d_ex expects rho <= 20\n");
    return -2;
  }
  if (do_ex == 1 && (nx1 > 3000 || ny1 > 3000 || nz1 > 3000)) {
    std::fprintf(stderr, "pmg_dminterp3d: nmax too small\n");
    return -2;
  }
  const int mi = (ni - 1) * irho + 1, mj = (nj - 1) * jrho + 1, mk = (nk - 1) * krho + 1;
  if (ord == 4 && do_ex != 1) {
    int rc = pmg_dmcopy3d(u1, u2, nx1, ny1, nz1, nx2, ny2, nz2, i1, j1, k1, i2, j2, k2, ni, nj, nk, si1, sj1, sk1,
                          irho, jrho, krho, stream);
    if (rc) return rc;
    size_t n = (size_t)mi * nj * nk;
    hipLaunchKernelGGL(k_interp4<0>, dim3(nblk(n)), dim3(NT), 0, stream, u2, nx2, ny2, nz2, i2, j2, k2, ni, nj, nk,
                       irho, jrho, krho);
    PMG_HOST_CHECK(hipGetLastError());
    if (nj == 1 && nk == 1) return 0;
    n = (size_t)mi * mj * nk;
    hipLaunchKernelGGL(k_interp4<1>, dim3(nblk(n)), dim3(NT), 0, stream, u2, nx2, ny2, nz2, i2, j2, k2, ni, nj, nk,
                       irho, jrho, krho);
    PMG_HOST_CHECK(hipGetLastError());
    if (nk == 1) return 0;
    n = (size_t)mi * mj * mk;
    hipLaunchKernelGGL(k_interp4<2>, dim3(nblk(n)), dim3(NT), 0, stream, u2, nx2, ny2, nz2, i2, j2, k2, ni, nj, nk,
                       irho, jrho, krho);
    PMG_HOST_CHECK(hipGetLastError());
    return 0;
  }
  if (do_ex != 1) {
    size_t n = (size_t)mi * mj * mk;
    hipLaunchKernelGGL(k_interp3d<double>, dim3(nblk(n)), dim3(NT), 0, stream, u1, u2, nx1,
                       ny1, nx2, ny2, i1, j1, k1, i2, j2, k2, ni, nj, nk, si1, sj1, sk1, irho,
                       jrho, krho);
    PMG_HOST_CHECK(hipGetLastError());
    return 0;
  }
  if (!chr1) return -1;
  auto kex_x = ord == 4 ? k_interp_ex_x<double, 4> : k_interp_ex_x<double, 2>;
  auto kex_y = ord == 4 ? k_interp_ex_y<double, 4> : k_interp_ex_y<double, 2>;
  auto kex_z = ord == 4 ? k_interp_ex_z<double, 4> : k_interp_ex_z<double, 2>;
  // common points, then the three line phases in place on u2
  int rc = pmg_dmcopy3d(u1, u2, nx1, ny1, nz1, nx2, ny2, nz2, i1, j1, k1, i2, j2, k2, ni, nj, nk,
                        si1, sj1, sk1, irho, jrho, krho, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(kex_x, dim3((nj * nk + 63) / 64), dim3(64), 0, stream, u2,
                     chr1, ex, nx1, ny1, nx2, ny2, i1, j1, k1, i2, j2, k2, ni, nj, nk, si1, sj1,
                     sk1, irho, jrho, krho);
  PMG_HOST_CHECK(hipGetLastError());
  if (nj == 1 && nk == 1) return 0;
  hipLaunchKernelGGL(kex_y, dim3((mi * nk + 63) / 64), dim3(64), 0, stream, u2,
                     chr1, ex, nx1, ny1, nx2, ny2, i1, j1, k1, i2, j2, k2, ni, nj, nk,
                     si1, sj1, sk1, irho, jrho, krho);
  PMG_HOST_CHECK(hipGetLastError());
  if (nk == 1) return 0;
  hipLaunchKernelGGL(kex_z, dim3((mi * mj + 63) / 64), dim3(64), 0, stream, u2,
                     chr1, ex, nx1, ny1, nz1, nx2, ny2, nz2, i1, j1, k1, i2, j2, k2,
                     ni, nj, nk, si1, sj1, sk1, irho, jrho, krho);
  PMG_HOST_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
