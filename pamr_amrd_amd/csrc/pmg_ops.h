// pmg_ops.h -- pointwise device operators shared by every kernel.
//
// The stencil expressions reproduce, operation by operation, the reference
// hooks and m_util_r8 routines cited at each function; with contraction
// disabled they give the same IEEE result as the reference's Fortran.
#pragma once
#include "pmg_common.h"

namespace pmg {

// x/d2, or x*inv when d2 is an exact power of two (then identical bits).
template <typename T, bool P2>
__device__ __forceinline__ T divd(T x, T d2, T inv) {
  if constexpr (P2) return x * inv; else return x / d2;
}

// a / b rounded to nearest from y = RN(1/b), without a division: q0 =
// RN(a y) is within 2 ulps of a/b; one correction with the exact remainder
// (an FMA), q1 = RN(q0 + (a - b q0) y), is faithful; a second one with a
// faithful quotient and a correctly rounded reciprocal returns RN(a/b)
// (P. Markstein, IA-64 and Elementary Functions, 2000, Thm. 8.10; J.-M.
// Muller et al., Handbook of Floating-Point Arithmetic, 2018, Sec. 4.7),
// absent underflow and overflow.  5 FP64 operations against the ~12 of the
// IEEE division sequence.
__device__ __forceinline__ double div_rn(double a, double b, double y) {
  const double q0 = a * y;
  const double r0 = __builtin_fma(-b, q0, a);
  const double q1 = __builtin_fma(r0, y, q0);
  const double r1 = __builtin_fma(-b, q1, a);
  return __builtin_fma(r1, y, q1);
}

// res / jac of the relaxation update V - res/Jac: with the plain operator
// in FP64, Jac is the level's constant and sp.ijac = RN(1/Jac), so div_rn
// gives the IEEE quotient's bits; otherwise the division itself.
template <typename T, int OP>
__device__ __forceinline__ T div_jac(T r, T jac, const Spacing<T> &sp) {
  if constexpr (OP == 0 && std::is_same<T, double>::value) return div_rn(r, jac, sp.ijac);
  else return r / jac;
}

// Plain Poisson operator L[V] = lap_h V - rho at one point
// (w/examples/nps/num_w.f:42-45; relax/residual terms at :136-140).
// Neighbour naming: e/w = i+1/i-1, n/s = j+1/j-1, t/b = k+1/k-1.
template <typename T, bool P2>
__device__ __forceinline__ T lop_poisson(T uc, T ue, T uw, T un, T us, T ut, T ub,
                                         T rho, const Spacing<T> &sp) {
  T lx = divd<T, P2>((ue - 2 * uc) + uw, sp.dx2, sp.idx2);
  T ly = divd<T, P2>((un - 2 * uc) + us, sp.dy2, sp.idy2);
  T lz = divd<T, P2>((ut - 2 * uc) + ub, sp.dz2, sp.idz2);
  return ((lx + ly) + lz) - rho;
}

// Variable-coefficient operator L[V] = div(a grad V) + c V - rho with
// arithmetic-mean face coefficients (project-defined hook in the nps style;
// the reference ships only constant-coefficient hooks).
template <typename T, bool P2>
__device__ __forceinline__ T lop_varcoef(T uc, T ue, T uw, T un, T us, T ut, T ub,
                                         T ac, T ae, T aw, T an, T as, T at, T ab,
                                         T cc, T rho, const Spacing<T> &sp, T *jac) {
  T fe = T(0.5) * (ac + ae), fw = T(0.5) * (ac + aw);
  T fn = T(0.5) * (ac + an), fs = T(0.5) * (ac + as);
  T ft = T(0.5) * (ac + at), fb = T(0.5) * (ac + ab);
  T lx = divd<T, P2>(fe * (ue - uc) - fw * (uc - uw), sp.dx2, sp.idx2);
  T ly = divd<T, P2>(fn * (un - uc) - fs * (uc - us), sp.dy2, sp.idy2);
  T lz = divd<T, P2>(ft * (ut - uc) - fb * (uc - ub), sp.dz2, sp.idz2);
  if (jac)
    *jac = ((divd<T, P2>(-(fe + fw), sp.dx2, sp.idx2) - divd<T, P2>(fn + fs, sp.dy2, sp.idy2)) -
            divd<T, P2>(ft + fb, sp.dz2, sp.idz2)) + cc;
  return (((lx + ly) + lz) + cc * uc) - rho;
}

// L at point p of a plain array (global memory neighbours).
template <typename T, int OP, bool P2>
__device__ __forceinline__ T lop_at(const T *V, const T *rho, const T *a, const T *c,
                                    size_t p, size_t sy, size_t sz, const Spacing<T> &sp,
                                    T *jac) {
  if constexpr (OP == 0) {
    if (jac) *jac = sp.jac;
    return lop_poisson<T, P2>(V[p], V[p + 1], V[p - 1], V[p + sy], V[p - sy], V[p + sz],
                              V[p - sz], rho[p], sp);
  } else {
    return lop_varcoef<T, P2>(V[p], V[p + 1], V[p - 1], V[p + sy], V[p - sy], V[p + sz],
                              V[p - sz], a[p], a[p + 1], a[p - 1], a[p + sy], a[p - sy],
                              a[p + sz], a[p - sz], c[p], rho[p], sp, jac);
  }
}

// ---------------------------------------------------------------------------
// Restriction point operators.  (i0,j0,k0) are 1-based fine indices as in
// src/m_util_r8.f; U(a,b,c) reads the fine array.
// ---------------------------------------------------------------------------
#define PMG_U1(a, b, c) u1[ix3((a)-1, (b)-1, (c)-1, nx1, ny1)]
#define PMG_CH(a, b, c) chr1[ix3((a)-1, (b)-1, (c)-1, nx1, ny1)]

// dmfwr3d, src/m_util_r8.f:775-823.  U(a, b, c) returns the fine value at
// the 1-based indices (a, b, c); the branch (full, plane, line or copy
// stencil) follows the position in the fine grid, as the Fortran does.
template <typename T, typename Acc>
__device__ __forceinline__ T fw_point_acc(const Acc &U, int nx1, int ny1, int nz1, int i0, int j0,
                                          int k0) {
  const int im = i0 - 1, jm = j0 - 1, km = k0 - 1, ip = i0 + 1, jp = j0 + 1, kp = k0 + 1;
  const bool xi = i0 > 1 && i0 <= nx1 - 1, yi = j0 > 1 && j0 <= ny1 - 1,
             zi = k0 > 1 && k0 <= nz1 - 1;
  if (xi && yi && zi) {
    return (U(im, jm, km) + 2 * U(i0, jm, km) + U(ip, jm, km) +
            2 * U(im, j0, km) + 4 * U(i0, j0, km) + 2 * U(ip, j0, km) +
            U(im, jp, km) + 2 * U(i0, jp, km) + U(ip, jp, km) +
            2 * U(im, jm, k0) + 4 * U(i0, jm, k0) + 2 * U(ip, jm, k0) +
            4 * U(im, j0, k0) + 8 * U(i0, j0, k0) + 4 * U(ip, j0, k0) +
            2 * U(im, jp, k0) + 4 * U(i0, jp, k0) + 2 * U(ip, jp, k0) +
            U(im, jm, kp) + 2 * U(i0, jm, kp) + U(ip, jm, kp) +
            2 * U(im, j0, kp) + 4 * U(i0, j0, kp) + 2 * U(ip, j0, kp) +
            U(im, jp, kp) + 2 * U(i0, jp, kp) + U(ip, jp, kp)) / T(64.0);
  } else if (xi && yi) {
    return (1 * U(im, jm, k0) + 2 * U(i0, jm, k0) + 1 * U(ip, jm, k0) +
            2 * U(im, j0, k0) + 4 * U(i0, j0, k0) + 2 * U(ip, j0, k0) +
            1 * U(im, jp, k0) + 2 * U(i0, jp, k0) + 1 * U(ip, jp, k0)) / T(16.0);
  } else if (yi && zi) {
    return (1 * U(i0, jm, km) + 2 * U(i0, jm, k0) + 1 * U(i0, jm, kp) +
            2 * U(i0, j0, km) + 4 * U(i0, j0, k0) + 2 * U(i0, j0, kp) +
            1 * U(i0, jp, km) + 2 * U(i0, jp, k0) + 1 * U(i0, jp, kp)) / T(16.0);
  } else if (xi && zi) {
    return (1 * U(im, j0, km) + 2 * U(im, j0, k0) + 1 * U(im, j0, kp) +
            2 * U(i0, j0, km) + 4 * U(i0, j0, k0) + 2 * U(i0, j0, kp) +
            1 * U(ip, j0, km) + 2 * U(ip, j0, k0) + 1 * U(ip, j0, kp)) / T(16.0);
  } else if (xi) {
    return (U(im, j0, k0) + 2 * U(i0, j0, k0) + U(ip, j0, k0)) / T(4.0);
  } else if (yi) {
    return (U(i0, jm, k0) + 2 * U(i0, j0, k0) + U(i0, jp, k0)) / T(4.0);
  } else if (zi) {
    return (U(i0, j0, km) + 2 * U(i0, j0, k0) + U(i0, j0, kp)) / T(4.0);
  }
  return U(i0, j0, k0);
}

template <typename T>
__device__ __forceinline__ T fw_point(const T *u1, int nx1, int ny1, int nz1, int i0, int j0,
                                      int k0) {
  auto U = [&](int a, int b, int c) -> T { return PMG_U1(a, b, c); };
  return fw_point_acc<T>(U, nx1, ny1, nz1, i0, j0, k0);
}

__device__ __forceinline__ int w3d(int a, int b, int c) {
  // w3d(3,3,3) of dmfwr3d_ex, src/m_util_r8.f:851-853: (2-|a-2|)(2-|b-2|)(2-|c-2|)
  return (2 - abs(a - 2)) * (2 - abs(b - 2)) * (2 - abs(c - 2)) / 1;
}
__device__ __forceinline__ int w3d_flat(int flat) {
  int a = flat % 3 + 1, b = (flat / 3) % 3 + 1, c = flat / 9 + 1;
  return w3d(a, b, c);
}
// 1D-k branch weight of dmfwr3d_ex (src/m_util_r8.f:931 reads w3d(i,1,kk),
// i the coarse loop index): table entry while (i-1)+9(kk-1) < 27, the
// intended w3d(1,1,kk) beyond it (there the reference read is undefined).
__device__ __forceinline__ int w3d_kbranch(int i, int kk) {
  int flat = (i - 1) + 9 * (kk - 1);
  return (flat >= 0 && flat < 27) ? w3d_flat(flat) : w3d(1, 1, kk);
}

// dmfwr3d_ex, src/m_util_r8.f:855-943.  iloop = coarse loop index i.
template <typename T>
__device__ __forceinline__ T fw_point_ex(const T *u1, const T *chr1, T ex, int nx1, int ny1,
                                         int nz1, int i0, int j0, int k0, int iloop) {
  const int im = i0 - 1, jm = j0 - 1, km = k0 - 1, ip = i0 + 1, jp = j0 + 1, kp = k0 + 1;
  const bool xi = i0 > 1 && i0 <= nx1 - 1, yi = j0 > 1 && j0 <= ny1 - 1,
             zi = k0 > 1 && k0 <= nz1 - 1;
  T sum = 0;
  int w = 0;
  if (xi && yi && zi) {
    for (int a = im; a <= ip; a++)
      for (int b = jm; b <= jp; b++)
        for (int c = km; c <= kp; c++)
          if (PMG_CH(a, b, c) != ex) {
            int ww = w3d(a - im + 1, b - jm + 1, c - km + 1);
            sum = sum + ww * PMG_U1(a, b, c);
            w += ww;
          }
  } else if (xi && yi) {
    for (int a = im; a <= ip; a++)
      for (int b = jm; b <= jp; b++)
        if (PMG_CH(a, b, k0) != ex) {
          int ww = w3d(a - im + 1, b - jm + 1, 1);
          sum = sum + ww * PMG_U1(a, b, k0);
          w += ww;
        }
  } else if (yi && zi) {
    for (int b = jm; b <= jp; b++)
      for (int c = km; c <= kp; c++)
        if (PMG_CH(i0, b, c) != ex) {
          int ww = w3d(1, b - jm + 1, c - km + 1);
          sum = sum + ww * PMG_U1(i0, b, c);
          w += ww;
        }
  } else if (xi && zi) {
    for (int a = im; a <= ip; a++)
      for (int c = km; c <= kp; c++)
        if (PMG_CH(a, j0, c) != ex) {
          int ww = w3d(a - im + 1, 1, c - km + 1);
          sum = sum + ww * PMG_U1(a, j0, c);
          w += ww;
        }
  } else if (xi) {
    for (int a = im; a <= ip; a++)
      if (PMG_CH(a, j0, k0) != ex) {
        int ww = w3d(a - im + 1, 1, 1);
        sum = sum + ww * PMG_U1(a, j0, k0);
        w += ww;
      }
  } else if (yi) {
    for (int b = jm; b <= jp; b++)
      if (PMG_CH(i0, b, k0) != ex) {
        int ww = w3d(1, b - jm + 1, 1);
        sum = sum + ww * PMG_U1(i0, b, k0);
        w += ww;
      }
  } else if (zi) {
    for (int c = km; c <= kp; c++)
      if (PMG_CH(i0, j0, c) != ex) {
        sum = sum + w3d(1, 1, c - km + 1) * PMG_U1(i0, j0, c);
        w += w3d_kbranch(iloop, c - km + 1);
      }
  } else {
    sum = PMG_U1(i0, j0, k0);
    w = 1;
  }
  if (w == 0 || PMG_CH(i0, j0, k0) == ex) return T(0);
  return sum / w;
}

// dmhwr3d, src/m_util_r8.f:675-717
template <typename T, bool EXC>
__device__ __forceinline__ T hw_point(const T *u1, const T *chr1, T ex, int nx1, int ny1,
                                      int nz1, int i0, int j0, int k0) {
  int w = 0;
  T sum = 0;
  if (i0 > 1 && i0 <= nx1 - 1) {
    if (!EXC || (PMG_CH(i0 - 1, j0, k0) != ex && PMG_CH(i0, j0, k0) != ex &&
                 PMG_CH(i0 + 1, j0, k0) != ex)) {
      w += 4;
      sum = sum + PMG_U1(i0 - 1, j0, k0) + 2 * PMG_U1(i0, j0, k0) + PMG_U1(i0 + 1, j0, k0);
    }
  }
  if (j0 > 1 && j0 <= ny1 - 1) {
    if (!EXC || (PMG_CH(i0, j0 - 1, k0) != ex && PMG_CH(i0, j0, k0) != ex &&
                 PMG_CH(i0, j0 + 1, k0) != ex)) {
      w += 4;
      sum = sum + PMG_U1(i0, j0 - 1, k0) + 2 * PMG_U1(i0, j0, k0) + PMG_U1(i0, j0 + 1, k0);
    }
  }
  if (k0 > 1 && k0 <= nz1 - 1) {
    if (!EXC || (PMG_CH(i0, j0, k0 - 1) != ex && PMG_CH(i0, j0, k0) != ex &&
                 PMG_CH(i0, j0, k0 + 1) != ex)) {
      w += 4;
      sum = sum + PMG_U1(i0, j0, k0 - 1) + 2 * PMG_U1(i0, j0, k0) + PMG_U1(i0, j0, k0 + 1);
    }
  }
  if (EXC && PMG_CH(i0, j0, k0) == ex) return T(0);
  if (w > 0) return sum / w;
  return PMG_U1(i0, j0, k0);
}
#undef PMG_U1
#undef PMG_CH

// One fine point of dvi4q1d (src/m_util_r8.f:65-137) on a line whose anchors
// are A(a), a = -1 .. n, at the 1-based line positions p0 + a*rho of a line
// of N points; f = a*rho + r (0 < r < rho) is the point's offset from the
// first anchor.  The first cell uses anchors 0..3 when p0 <= rho (no anchor
// to its left), the last cell anchors n-4..n-1 when p0 + (n+2)*rho > N,
// every other cell a-1..a+2; n < 4 falls back to dvi2q1d (:332-360).  The
// literals are the reference's 0.1666 6666 6666 6667 d0 (16 digits: fixed-
// form blanks are not significant) and the operations keep its order.
template <typename F>
__device__ __forceinline__ double vi4_point(int n, int rho, int p0, int N, int f, const F &A) {
  const int a = f / rho, r = f % rho;
  if (n < 4) {
    const double sg = (r * 1.0) / rho;
    const double c0 = 1 - sg, c1 = sg;
    return c0 * A(a) + c1 * A(a + 1);
  }
  const double half = 0.5, mhalf = -0.5, sixth = 0.1666666666666667, msixth = -0.1666666666666667;
  double sg;
  int b;
  if (p0 <= rho && a == 0) {
    sg = (r * 1.0) / rho;
    b = 0;
  } else if (p0 + (n + 2) * rho > N && a == n - 2) {
    sg = 2.0 + (r * 1.0) / rho;
    b = n - 4;
  } else {
    sg = 1.0 + (r * 1.0) / rho;
    b = a - 1;
  }
  const double c0 = msixth * (sg - 1.0) * (sg - 2.0) * (sg - 3.0);
  const double c1 = half * sg * (sg - 2.0) * (sg - 3.0);
  const double c2 = mhalf * sg * (sg - 1.0) * (sg - 3.0);
  const double c3 = sixth * sg * (sg - 1.0) * (sg - 2.0);
  return c0 * A(b) + c1 * A(b + 1) + c2 * A(b + 2) + c3 * A(b + 3);
}

}  // namespace pmg
