/*
 * mg_oracle.c -- CPU oracle for the PAMR/AMRD multigrid hot path.
 *
 * TEST INFRASTRUCTURE ONLY: called by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker and as the timed CPU baseline
 * ("port" of the reference CPU path).  The product library never links it.
 *
 * What it restates (reference files under lorenzorossi7/pamr_amrd):
 *   src/m_util_r8.f:20-50    dmcopy3d            -> or_dmcopy3d
 *   src/m_util_r8.f:332-360  dvi2q1d             -> or_dvi2q1d
 *   src/m_util_r8.f:368-421  dvi2q1d_ex          -> or_dvi2q1d_ex
 *   src/m_util_r8.f:426-456  init_chrt/and_chrt  -> chrt_init/chrt_and
 *   src/m_util_r8.f:65-137   dvi4q1d             -> or_dvi4q1d
 *   src/m_util_r8.f:146-319  dvi4q1d_ex          -> or_dvi4q1d_ex
 *   src/m_util_r8.f:473-637  dminterp3d (ord 2, 4) -> or_dminterp3d
 *   src/m_util_r8.f:649-723  dmhwr3d             -> or_dmhwr3d
 *   src/m_util_r8.f:733-949  dmfwr3d(_ex)        -> or_dmfwr3d
 *   w/examples/nps/num_w.f:18-157 lop/residual/relax (plain Poisson hooks,
 *                             vertex-centred point counts) -> or_lop/...
 *   amrd/mg.c:20-201,242-355 solve_elliptics / vcycle / clear_MG_vars /
 *                             compute_MG_rhs / compute_cgc -> or_mg_*
 *   amrd/mg.c:362-488        set_cmask_child / set_cmask_bdy
 *   src/gh.c:1222-1365,1427-1467 build_mgh coarsening, calc_island_no
 *   src/transfer.c:73-95,946-1163 calc_gs_ibbox, local_transfer index math
 *   amrd/util.c:73-103       call_rr_app (MAX over grids; one grid here)
 *
 * Pinning (see tests/test_oracle_*.py): the m_util_r8 and hook restatements
 * are compared bit-for-bit with the reference Fortran compiled from source
 * into oracle/_ref (oracle/Makefile); the driver is pinned by closed forms
 * (discrete eigenfunction solve, FAS == correction scheme for a linear L,
 * zero-source fixed point, affine exactness of the transfers) and by an
 * independent composite-FAS formulation built from the reference kernels.
 * The variable-coefficient operator (op=1) and the network oracle have no
 * reference counterpart; they are pinned by closed forms (constant a reduces
 * to the Poisson hook, finite-difference gradient checks).
 */
#include "mg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAX(a, b) ((a) > (b) ? (a) : (b))
#define MIN(a, b) ((a) < (b) ? (a) : (b))
/* 1-based Fortran element (i,j,k) of an nx x ny x nz array */
#define F3(a, i, j, k, nx, ny) \
  (a)[((size_t)(i)-1) + (size_t)(nx) * (((size_t)(j)-1) + (size_t)(ny) * ((size_t)(k)-1))]

/* ======================================================================== */
/* m_util_r8 restatements                                                    */
/* ======================================================================== */

/* src/m_util_r8.f:20-50 */
void or_dmcopy3d(const double *u1, double *u2, int nx1, int ny1, int nz1,
                 int nx2, int ny2, int nz2, int i1, int j1, int k1,
                 int i2, int j2, int k2, int ni, int nj, int nk,
                 int si1, int sj1, int sk1, int si2, int sj2, int sk2)
{
  int i, j, k;
  (void)nz1; (void)nz2;
  if (i1 < 1 || (i1 + (ni - 1) * si1) > nx1) return;
  if (i2 < 1 || (i2 + (ni - 1) * si2) > nx2) return;
  if (j1 < 1 || (j1 + (nj - 1) * sj1) > ny1) return;
  if (j2 < 1 || (j2 + (nj - 1) * sj2) > ny2) return;
  if (k1 < 1 || (k1 + (nk - 1) * sk1) > nz1) return;
  if (k2 < 1 || (k2 + (nk - 1) * sk2) > nz2) return;
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++)
      for (k = 1; k <= nk; k++)
        F3(u2, i2 + (i - 1) * si2, j2 + (j - 1) * sj2, k2 + (k - 1) * sk2, nx2, ny2) =
            F3(u1, i1 + (i - 1) * si1, j1 + (j - 1) * sj1, k1 + (k - 1) * sk1, nx1, ny1);
}

/* src/m_util_r8.f:332-360.  v1/v2 point at element 1 of the (strided) line;
 * they may alias (dminterp3d interpolates in place). */
void or_dvi2q1d(double *v1, double *v2, int i1, int i2, int n, int irho,
                int s1, int s2)
{
  int i, j;
  double c0, c1, sg;
  for (i = 1; i <= n; i++) v2[irho * (i - 1) * s2 + i2 - 1] = v1[(i - 1) * s1 + i1 - 1];
  for (j = 1; j <= irho - 1; j++) {
    sg = (j * 1.0) / irho;
    c0 = 1 - sg;
    c1 = sg;
    for (i = 0; i <= n - 2; i++)
      v2[(irho * i + j) * s2 + i2 - 1] = c0 * v1[i * s1 + i1 - 1] + c1 * v1[(i + 1) * s1 + i1 - 1];
  }
}

/* src/m_util_r8.f:65-137: cubic (4-point Lagrange) irho:1 interpolation;
 * n1 is the length of v1's array (the right-boundary test), n < 4 falls
 * back to dvi2q1d.  Interior cells take the anchors i-1..i+2 (also outside
 * the segment when they exist), the first / last cell the four anchors
 * inside when no anchor lies beyond. */
void or_dvi4q1d(double *v1, double *v2, int i1, int i2, int n1, int n, int irho, int s1, int s2)
{
  int i, j, is, ie;
  double c0, c1, c2, c3, sg;
  /* the reference's literals are 0.1666 6666 6666 6667 d0: fixed-form blanks
   * are not significant, so 16 digits (a different double than 1/6's) */
  const double half = 0.5, mhalf = -0.5, sixth = 0.1666666666666667, msixth = -0.1666666666666667;
#define V1(q) v1[(q)-1]
#define V2(q) v2[(q)-1]
  if (n < 4) {
    or_dvi2q1d(v1, v2, i1, i2, n, irho, s1, s2);
    return;
  }
  for (i = 1; i <= n; i++) V2(irho * (i - 1) * s2 + i2) = V1((i - 1) * s1 + i1);
  if (i1 <= s1) {
    for (j = 1; j <= irho - 1; j++) {
      sg = (j * 1.0) / irho;
      c0 = msixth * (sg - 1.0) * (sg - 2.0) * (sg - 3.0);
      c1 = half * sg * (sg - 2.0) * (sg - 3.0);
      c2 = mhalf * sg * (sg - 1.0) * (sg - 3.0);
      c3 = sixth * sg * (sg - 1.0) * (sg - 2.0);
      V2(j * s2 + i2) = c0 * V1(i1) + c1 * V1(s1 + i1) + c2 * V1(2 * s1 + i1) + c3 * V1(3 * s1 + i1);
    }
    is = 1;
  } else {
    is = 0;
  }
  if ((i1 + n * s1) > (n1 - 2 * s1)) {
    ie = n - 3;
    for (j = 1; j <= irho - 1; j++) {
      sg = 2.0 + (j * 1.0) / irho;
      c0 = msixth * (sg - 1.0) * (sg - 2.0) * (sg - 3.0);
      c1 = half * sg * (sg - 2.0) * (sg - 3.0);
      c2 = mhalf * sg * (sg - 1.0) * (sg - 3.0);
      c3 = sixth * sg * (sg - 1.0) * (sg - 2.0);
      V2((irho * (n - 2) + j) * s2 + i2) = c0 * V1((n - 4) * s1 + i1) + c1 * V1((n - 3) * s1 + i1) +
                                          c2 * V1((n - 2) * s1 + i1) + c3 * V1((n - 1) * s1 + i1);
    }
  } else {
    ie = n - 2;
  }
  for (j = 1; j <= irho - 1; j++) {
    sg = 1.0 + (j * 1.0) / irho;
    c0 = msixth * (sg - 1.0) * (sg - 2.0) * (sg - 3.0);
    c1 = half * sg * (sg - 2.0) * (sg - 3.0);
    c2 = mhalf * sg * (sg - 1.0) * (sg - 3.0);
    c3 = sixth * sg * (sg - 1.0) * (sg - 2.0);
    for (i = is; i <= ie; i++)
      V2((irho * i + j) * s2 + i2) = c0 * V1((i - 1) * s1 + i1) + c1 * V1(i * s1 + i1) + c2 * V1((i + 1) * s1 + i1) +
                                     c3 * V1((i + 2) * s1 + i1);
  }
#undef V1
#undef V2
}

/* src/m_util_r8.f:146-319: dvi4q1d with excision.  Extrapolates excised
 * anchors next to non-excised runs (in place in v1), copies the common
 * points, then writes the first / last / interior cubic cells whose four
 * anchors are all non-excised; a non-excised run shorter than 4 anchors
 * first sends the whole line through dvi2q1d_ex.  chr1/i3/s3 index the
 * coarse mask as in or_dvi2q1d_ex.  The conditions and the '(i1s-4*s1)'
 * bound of the right-side extrapolation are kept as written. */
void or_dvi4q1d_ex(double *v1, double *v2, int i1, int i2, int n1, int n, int irho, int s1, int s2,
                   const double *chr1, int i3, int s3, double ex)
{
  double c0[20], c1[20], c2[20], c3[20], sg;
  int i, j, cn, j0, i1s, i1e, i10, i30, init_2;
  const double half = 0.5, mhalf = -0.5, sixth = 0.1666666666666667, msixth = -0.1666666666666667;
#define V1(q) v1[(q)-1]
#define V2(q) v2[(q)-1]
#define C1(q) chr1[(q)-1]
  if (irho > 20) {
    fprintf(stderr, "or_dvi4q1d_ex: irho expected to be less than 20\n");
    abort();
  }
  i = 3;
  i1s = i1;
  while (i > 0 && i1s > s1) {
    i1s -= s1;
    i--;
  }
  i = 3;
  i1e = i1 + (n - 1) * s1;
  while (i > 0 && i1e < (n1 - s1)) {
    i1e += s1;
    i--;
  }
  cn = 0;
  init_2 = 0;
  for (i = i1s; i <= i1e; i += s1) {
    if (C1((i - i1) * s3 / s1 + i3) == ex || i == i1e) {
      if (cn > 0 && cn < 4)
        init_2 = 1;
      else
        cn = 0;
    } else {
      cn++;
    }
  }
  if (init_2) or_dvi2q1d_ex(v1, v2, i1, i2, n1, n, irho, s1, s2, chr1, i3, s3, ex);
  for (i = i1s; i <= i1e; i += s1) {
    if (C1(i3 + s3 * (i - i1) / s1) == ex) {
      if (i >= (i1s + 3 * s1) && C1(i3 + s3 * (i - s1 - i1) / s1) != ex && C1(i3 + s3 * (i - 2 * s1 - i1) / s1) != ex &&
          C1(i3 + s3 * (i - 3 * s1 - i1) / s1) != ex) {
        if (i >= (i1s + 4 * s1) && C1(i3 + s3 * (i - 4 * s1 - i1) / s1) != ex)
          V1(i) = 4 * (V1(i - s1) + V1(i - 3 * s1)) - 6 * V1(i - 2 * s1) - V1(i - 4 * s1);
        else
          V1(i) = 3 * (V1(i - s1) - V1(i - 2 * s1)) + V1(i - 3 * s1);
      } else if (i <= (i1e - 3 * s1) && C1(i3 + s3 * (i + s1 - i1) / s1) != ex &&
                 C1(i3 + s3 * (i + 2 * s1 - i1) / s1) != ex && C1(i3 + s3 * (i + 3 * s1 - i1) / s1) != ex) {
        if (i <= (i1s - 4 * s1) && C1(i3 + s3 * (i + 4 * s1 - i1) / s1) != ex)
          V1(i) = 4 * (V1(i + s1) + V1(i + 3 * s1)) - 6 * V1(i + 2 * s1) - V1(i + 4 * s1);
        else
          V1(i) = 3 * (V1(i + s1) - V1(i + 2 * s1)) + V1(i + 3 * s1);
      }
    }
    if (i >= i1 && i <= (i1 + (n - 1) * s1)) V2(irho * ((i - i1) / s1) * s2 + i2) = V1(i);
  }
  /* first cells of non-excised runs */
  for (j = 1; j <= irho - 1; j++) {
    sg = (j * 1.0) / irho;
    c0[j - 1] = msixth * (sg - 1.0) * (sg - 2.0) * (sg - 3.0);
    c1[j - 1] = half * sg * (sg - 2.0) * (sg - 3.0);
    c2[j - 1] = mhalf * sg * (sg - 1.0) * (sg - 3.0);
    c3[j - 1] = sixth * sg * (sg - 1.0) * (sg - 2.0);
  }
  for (i = 1; i <= n - 1; i++) {
    i10 = (i - 1) * s1 + i1;
    i30 = (i - 1) * s3 + i3;
    j0 = (i - 1) * irho * s2 + i2;
    if (i10 <= (i1e - 3 * s1) && C1(i30 + s3) != ex && C1(i30 + 2 * s3) != ex &&
        ((i == 1 && i10 == i1s) || (i10 >= (i1s + s1) && C1(i30) != ex && C1(i30 - s3) == ex) || C1(i30) == ex)) {
      for (j = 1; j <= irho - 1; j++)
        V2(j * s2 + j0) = c0[j - 1] * V1(i10) + c1[j - 1] * V1(s1 + i10) + c2[j - 1] * V1(2 * s1 + i10) +
                          c3[j - 1] * V1(3 * s1 + i10);
    }
  }
  /* last cells of non-excised runs */
  for (j = 1; j <= irho - 1; j++) {
    sg = 2.0 + (j * 1.0) / irho;
    c0[j - 1] = msixth * (sg - 1.0) * (sg - 2.0) * (sg - 3.0);
    c1[j - 1] = half * sg * (sg - 2.0) * (sg - 3.0);
    c2[j - 1] = mhalf * sg * (sg - 1.0) * (sg - 3.0);
    c3[j - 1] = sixth * sg * (sg - 1.0) * (sg - 2.0);
  }
  for (i = 1; i <= n - 1; i++) {
    i10 = (i - 1) * s1 + i1;
    i30 = (i - 1) * s3 + i3;
    j0 = (i - 1) * irho * s2 + i2;
    if (i10 >= (i1s + 2 * s1) && C1(i30 - s3) != ex &&
        ((i == (n - 1) && i10 == (i1e - s1)) || (i10 <= (i1e - s1) && C1(i30) != ex && C1(i30 + s3) == ex))) {
      for (j = 1; j <= irho - 1; j++)
        V2(j * s2 + j0) = c0[j - 1] * V1(i10 - 2 * s1) + c1[j - 1] * V1(i10 - s1) + c2[j - 1] * V1(i10) +
                          c3[j - 1] * V1(i10 + s1);
    }
  }
  /* interior cells */
  for (j = 1; j <= irho - 1; j++) {
    sg = 1.0 + (j * 1.0) / irho;
    c0[j - 1] = msixth * (sg - 1.0) * (sg - 2.0) * (sg - 3.0);
    c1[j - 1] = half * sg * (sg - 2.0) * (sg - 3.0);
    c2[j - 1] = mhalf * sg * (sg - 1.0) * (sg - 3.0);
    c3[j - 1] = sixth * sg * (sg - 1.0) * (sg - 2.0);
  }
  for (i = 1; i <= n; i++) {
    i10 = (i - 1) * s1 + i1;
    i30 = (i - 1) * s3 + i3;
    j0 = (i - 1) * irho * s2 + i2;
    if (C1(i30) != ex && i10 >= (i1s + s1) && C1(i30 - s3) != ex && i10 <= (i1e - 2 * s1) && C1(i30 + s3) != ex) {
      for (j = 1; j <= irho - 1; j++)
        V2(j0 + j * s2) = c0[j - 1] * V1(i10 - s1) + c1[j - 1] * V1(i10) + c2[j - 1] * V1(i10 + s1) +
                          c3[j - 1] * V1(i10 + 2 * s1);
    }
  }
#undef V1
#undef V2
#undef C1
}

/* src/m_util_r8.f:368-421 (extrapolates excised anchors of v1, in place) */
void or_dvi2q1d_ex(double *v1, double *v2, int i1, int i2, int n1, int n,
                   int irho, int s1, int s2, const double *chr1, int i3,
                   int s3, double ex)
{
  int i, j, i10, i30;
  double c0, c1, sg;
#define V1(q) v1[(q)-1]
#define C1(q) chr1[(q)-1]
  for (i = 1; i <= n; i++) {
    i10 = (i - 1) * s1 + i1;
    i30 = (i - 1) * s3 + i3;
    if (C1(i30) == ex) {
      if (i10 > 3 * s1 && C1(i30 - s3) != ex && C1(i30 - 2 * s3) != ex && C1(i30 - 3 * s3) != ex)
        V1(i10) = 3 * (V1(i10 - s1) - V1(i10 - 2 * s1)) + V1(i10 - 3 * s1);
      else if (i10 <= (n1 - 3 * s1) && C1(i30 + s3) != ex && C1(i30 + 2 * s3) != ex &&
               C1(i30 + 3 * s3) != ex)
        V1(i10) = 3 * (V1(i10 + s1) - V1(i10 + 2 * s1)) + V1(i10 + 3 * s1);
      else if (i10 > 2 * s1 && C1(i30 - s3) != ex && C1(i30 - 2 * s3) != ex)
        V1(i10) = 2 * V1(i10 - s1) - V1(i10 - 2 * s1);
      else if (i10 <= (n1 - 2 * s1) && C1(i30 + s3) != ex && C1(i30 + 2 * s3) != ex)
        V1(i10) = 2 * V1(i10 + s1) - V1(i10 + 2 * s1);
      else if (i10 > s1 && C1(i30 - s3) != ex)
        V1(i10) = V1(i10 - s1);
      else if (i10 <= (n1 - s1) && C1(i30 + s3) != ex)
        V1(i10) = V1(i10 + s1);
    }
    v2[irho * (i - 1) * s2 + i2 - 1] = V1(i10);
  }
  for (j = 1; j <= irho - 1; j++) {
    sg = (j * 1.0) / irho;
    c0 = 1 - sg;
    c1 = sg;
    for (i = 0; i <= n - 2; i++) {
      i10 = (irho * i + j) * s2 + i2;
      v2[i10 - 1] = c0 * V1(i * s1 + i1) + c1 * V1((i + 1) * s1 + i1);
    }
  }
#undef V1
#undef C1
}

/* src/m_util_r8.f:426-456 */
static void chrt_init(double *chrt, const double *chr1, int n, int s3)
{
  int i;
  for (i = 1; i <= n; i++) chrt[i - 1] = chr1[(i - 1) * s3];
}
static void chrt_and(double *chrt, const double *chr1, int n, int s3, double ex)
{
  int i;
  for (i = 1; i <= n; i++) {
    if (chrt[i - 1] == ex && chr1[(i - 1) * s3] == ex)
      chrt[i - 1] = ex;
    else
      chrt[i - 1] = ex + 1;
  }
}

/* src/m_util_r8.f:473-637: ord 2 and 4, with or without excision */
void or_dminterp3d(double *u1, double *u2, int nx1, int ny1, int nz1,
                   int nx2, int ny2, int nz2, int i1, int j1, int k1,
                   int i2, int j2, int k2, int ni, int nj, int nk,
                   int si1, int sj1, int sk1, int irho, int jrho, int krho,
                   int ord, const double *chr1, double ex, int do_ex)
{
  int i, j, k;
  const int nmax = 3000;
  double chrt[3000];
  if (i1 < 1 || (i1 + ni - 1) * si1 > nx1) return;
  if (i2 < 1 || (i2 + (ni - 1) * irho) > nx2) return;
  if (j1 < 1 || (j1 + nj - 1) * sj1 > ny1) return;
  if (j2 < 1 || (j2 + (nj - 1) * jrho) > ny2) return;
  if (k1 < 1 || (k1 + nk - 1) * sk1 > nz1) return;
  if (k2 < 1 || (k2 + (nk - 1) * krho) > nz2) return;
  if (!(ord == 2 || ord == 4)) {
    fprintf(stderr, "or_dminterp3d: only ord=2,4 currently supported\n");
    abort();
  }
  if (do_ex == 1 && (nx1 > nmax || ny1 > nmax || nz1 > nmax)) {
    fprintf(stderr, "or_dminterp3d: nmax too small\n");
    abort();
  }
  or_dmcopy3d(u1, u2, nx1, ny1, nz1, nx2, ny2, nz2, i1, j1, k1, i2, j2, k2, ni, nj, nk,
              si1, sj1, sk1, irho, jrho, krho);
  /* x lines through anchor (j,k) */
  for (j = j2; j <= j2 + (nj - 1) * jrho; j += jrho)
    for (k = k2; k <= k2 + (nk - 1) * krho; k += krho) {
      double *line = &F3(u2, 1, j, k, nx2, ny2);
      if (do_ex == 1) {
        const double *cl = &F3(chr1, 1, j1 + sj1 * (j - j2) / jrho, k1 + sk1 * (k - k2) / krho, nx1, ny1);
        if (ord == 4)
          or_dvi4q1d_ex(line, line, i2, i2, nx2, ni, irho, irho, 1, cl, i1, si1, ex);
        else
          or_dvi2q1d_ex(line, line, i2, i2, nx2, ni, irho, irho, 1, cl, i1, si1, ex);
      } else if (ord == 4) {
        or_dvi4q1d(line, line, i2, i2, nx2, ni, irho, irho, 1);
      } else {
        or_dvi2q1d(line, line, i2, i2, ni, irho, irho, 1);
      }
    }
  if (nj == 1 && nk == 1) return;
  /* y lines for every i, anchor k */
  for (i = i2; i <= i2 + (ni - 1) * irho; i++)
    for (k = k2; k <= k2 + (nk - 1) * krho; k += krho) {
      double *col = &F3(u2, i, 1, k, nx2, ny2);
      if (do_ex == 1) {
        chrt_init(chrt, &F3(chr1, i1 + si1 * ((i - i2) / irho), 1, k1 + sk1 * ((k - k2) / krho), nx1, ny1), ny1, nx1);
        if ((i - i2) % irho != 0)
          chrt_and(chrt, &F3(chr1, i1 + si1 * ((i - i2) / irho + 1), 1, k1 + sk1 * ((k - k2) / krho), nx1, ny1), ny1, nx1, ex);
        if (ord == 4)
          or_dvi4q1d_ex(col, col, 1 + (j2 - 1) * nx2, 1 + (j2 - 1) * nx2, nx2 * (ny2 - 1) + 1, nj, jrho,
                        jrho * nx2, nx2, chrt, j1, sj1, ex);
        else
          or_dvi2q1d_ex(col, col, 1 + (j2 - 1) * nx2, 1 + (j2 - 1) * nx2, nx2 * (ny2 - 1) + 1, nj, jrho,
                        jrho * nx2, nx2, chrt, j1, sj1, ex);
      } else if (ord == 4) {
        or_dvi4q1d(col, col, 1 + (j2 - 1) * nx2, 1 + (j2 - 1) * nx2, nx2 * (ny2 - 1) + 1, nj, jrho, jrho * nx2, nx2);
      } else {
        or_dvi2q1d(col, col, 1 + (j2 - 1) * nx2, 1 + (j2 - 1) * nx2, nj, jrho, jrho * nx2, nx2);
      }
    }
  if (nk == 1) return;
  /* z lines for every (i,j) */
  for (i = i2; i <= i2 + (ni - 1) * irho; i++)
    for (j = j2; j <= j2 + (nj - 1) * jrho; j++) {
      double *col = &F3(u2, i, j, 1, nx2, ny2);
      if (do_ex == 1) {
        int ic = i1 + si1 * ((i - i2) / irho), jc = j1 + sj1 * ((j - j2) / jrho);
        int nxy = nx1 * ny1;
        chrt_init(chrt, &F3(chr1, ic, jc, 1, nx1, ny1), nz1, nxy);
        if ((i - i2) % irho != 0)
          chrt_and(chrt, &F3(chr1, i1 + si1 * ((i - i2) / irho + 1), jc, 1, nx1, ny1), nz1, nxy, ex);
        if ((j - j2) % jrho != 0)
          chrt_and(chrt, &F3(chr1, ic, j1 + sj1 * ((j - j2) / jrho + 1), 1, nx1, ny1), nz1, nxy, ex);
        if ((j - j2) % jrho != 0 && (i - i2) % irho != 0)
          chrt_and(chrt, &F3(chr1, i1 + si1 * ((i - i2) / irho + 1), j1 + sj1 * ((j - j2) / jrho + 1), 1, nx1, ny1),
                   nz1, nxy, ex);
        if (ord == 4)
          or_dvi4q1d_ex(col, col, (k2 - 1) * nx2 * ny2 + 1, (k2 - 1) * nx2 * ny2 + 1,
                        nx2 * ny2 * (nz2 - 1) + 1, nk, krho, krho * nx2 * ny2, nx2 * ny2, chrt, k1, sk1, ex);
        else
          or_dvi2q1d_ex(col, col, (k2 - 1) * nx2 * ny2 + 1, (k2 - 1) * nx2 * ny2 + 1,
                        nx2 * ny2 * (nz2 - 1) + 1, nk, krho, krho * nx2 * ny2, nx2 * ny2, chrt, k1, sk1, ex);
      } else if (ord == 4) {
        or_dvi4q1d(col, col, (k2 - 1) * nx2 * ny2 + 1, (k2 - 1) * nx2 * ny2 + 1, nx2 * ny2 * (nz2 - 1) + 1, nk,
                   krho, krho * nx2 * ny2, nx2 * ny2);
      } else {
        or_dvi2q1d(col, col, (k2 - 1) * nx2 * ny2 + 1, (k2 - 1) * nx2 * ny2 + 1, nk, krho,
                   krho * nx2 * ny2, nx2 * ny2);
      }
    }
}

/* src/m_util_r8.f:649-723 */
void or_dmhwr3d(const double *u1, double *u2, int nx1, int ny1, int nz1,
                int nx2, int ny2, int nz2, int i1, int j1, int k1,
                int i2, int j2, int k2, int ni, int nj, int nk,
                int si1, int sj1, int sk1, int si2, int sj2, int sk2,
                const double *chr1, double ex, int do_ex)
{
  int i, j, k, w, i0, j0, k0;
  double sum;
  if (i1 < 1 || (i1 + (ni - 1) * si1) > nx1) return;
  if (i2 < 1 || (i2 + (ni - 1) * si2) > nx2) return;
  if (j1 < 1 || (j1 + (nj - 1) * sj1) > ny1) return;
  if (j2 < 1 || (j2 + (nj - 1) * sj2) > ny2) return;
  if (k1 < 1 || (k1 + (nk - 1) * sk1) > nz1) return;
  if (k2 < 1 || (k2 + (nk - 1) * sk2) > nz2) return;
#define U1(a, b, c) F3(u1, a, b, c, nx1, ny1)
#define CH(a, b, c) F3(chr1, a, b, c, nx1, ny1)
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++)
      for (k = 1; k <= nk; k++) {
        double *d = &F3(u2, i2 + (i - 1) * si2, j2 + (j - 1) * sj2, k2 + (k - 1) * sk2, nx2, ny2);
        i0 = i1 + (i - 1) * si1;
        j0 = j1 + (j - 1) * sj1;
        k0 = k1 + (k - 1) * sk1;
        w = 0;
        sum = 0;
        if (i0 > 1 && i0 <= nx1 - 1) {
          if (do_ex == 0 || (CH(i0 - 1, j0, k0) != ex && CH(i0, j0, k0) != ex && CH(i0 + 1, j0, k0) != ex)) {
            w += 4;
            sum = sum + U1(i0 - 1, j0, k0) + 2 * U1(i0, j0, k0) + U1(i0 + 1, j0, k0);
          }
        }
        if (j0 > 1 && j0 <= ny1 - 1) {
          if (do_ex == 0 || (CH(i0, j0 - 1, k0) != ex && CH(i0, j0, k0) != ex && CH(i0, j0 + 1, k0) != ex)) {
            w += 4;
            sum = sum + U1(i0, j0 - 1, k0) + 2 * U1(i0, j0, k0) + U1(i0, j0 + 1, k0);
          }
        }
        if (k0 > 1 && k0 <= nz1 - 1) {
          if (do_ex == 0 || (CH(i0, j0, k0 - 1) != ex && CH(i0, j0, k0) != ex && CH(i0, j0, k0 + 1) != ex)) {
            w += 4;
            sum = sum + U1(i0, j0, k0 - 1) + 2 * U1(i0, j0, k0) + U1(i0, j0, k0 + 1);
          }
        }
        if (do_ex == 1 && CH(i0, j0, k0) == ex)
          *d = 0;
        else if (w > 0)
          *d = sum / w;
        else
          *d = U1(i0, j0, k0);
      }
#undef U1
#undef CH
}

/* w3d weight table of dmfwr3d_ex, Fortran order w3d(3,3,3) */
static const int w3d_tab[27] = {1, 2, 1, 2, 4, 2, 1, 2, 1, 2, 4, 2, 4, 8, 4, 2, 4, 2,
                                1, 2, 1, 2, 4, 2, 1, 2, 1};
#define W3D(a, b, c) w3d_tab[((a)-1) + 3 * ((b)-1) + 9 * ((c)-1)]

/* The 1D-k branch of dmfwr3d_ex accumulates w3d(i,1,k01-km+1), i being the
 * coarse loop index (src/m_util_r8.f:931).  For flat index (i-1)+9*(kk-1) <
 * 27 the reference reads that table entry and so do we; beyond the table the
 * reference read is out of bounds (undefined) and we use the intended
 * w3d(1,1,kk).  The product library follows the same rule. */
static int w3d_kbranch(int i, int kk)
{
  int flat = (i - 1) + 9 * (kk - 1);
  if (flat >= 0 && flat < 27) return w3d_tab[flat];
  return W3D(1, 1, kk);
}

static void dmfwr3d_ex(const double *u1, double *u2, int nx1, int ny1, int nz1,
                       int nx2, int ny2, int nz2, int i1, int j1, int k1,
                       int i2, int j2, int k2, int ni, int nj, int nk,
                       int si1, int sj1, int sk1, int si2, int sj2, int sk2,
                       const double *chr1, double ex)
{
  int i, j, k, i0, j0, k0, im, jm, km, ip, jp, kp, i01, j01, k01, w;
  double sum;
  (void)nz1; (void)nz2;
#define U1(a, b, c) F3(u1, a, b, c, nx1, ny1)
#define CH(a, b, c) F3(chr1, a, b, c, nx1, ny1)
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++)
      for (k = 1; k <= nk; k++) {
        i0 = i1 + (i - 1) * si1; j0 = j1 + (j - 1) * sj1; k0 = k1 + (k - 1) * sk1;
        im = i0 - 1; jm = j0 - 1; km = k0 - 1; ip = i0 + 1; jp = j0 + 1; kp = k0 + 1;
        sum = 0;
        w = 0;
        if (i0 > 1 && i0 <= nx1 - 1 && j0 > 1 && j0 <= ny1 - 1 && k0 > 1 && k0 <= nz1 - 1) {
          for (i01 = im; i01 <= ip; i01++)
            for (j01 = jm; j01 <= jp; j01++)
              for (k01 = km; k01 <= kp; k01++)
                if (CH(i01, j01, k01) != ex) {
                  sum = sum + W3D(i01 - im + 1, j01 - jm + 1, k01 - km + 1) * U1(i01, j01, k01);
                  w = w + W3D(i01 - im + 1, j01 - jm + 1, k01 - km + 1);
                }
        } else if (i0 > 1 && i0 <= nx1 - 1 && j0 > 1 && j0 <= ny1 - 1) {
          for (i01 = im; i01 <= ip; i01++)
            for (j01 = jm; j01 <= jp; j01++)
              if (CH(i01, j01, k0) != ex) {
                sum = sum + W3D(i01 - im + 1, j01 - jm + 1, 1) * U1(i01, j01, k0);
                w = w + W3D(i01 - im + 1, j01 - jm + 1, 1);
              }
        } else if (j0 > 1 && j0 <= ny1 - 1 && k0 > 1 && k0 <= nz1 - 1) {
          for (j01 = jm; j01 <= jp; j01++)
            for (k01 = km; k01 <= kp; k01++)
              if (CH(i0, j01, k01) != ex) {
                sum = sum + W3D(1, j01 - jm + 1, k01 - km + 1) * U1(i0, j01, k01);
                w = w + W3D(1, j01 - jm + 1, k01 - km + 1);
              }
        } else if (i0 > 1 && i0 <= nx1 - 1 && k0 > 1 && k0 <= nz1 - 1) {
          for (i01 = im; i01 <= ip; i01++)
            for (k01 = km; k01 <= kp; k01++)
              if (CH(i01, j0, k01) != ex) {
                sum = sum + W3D(i01 - im + 1, 1, k01 - km + 1) * U1(i01, j0, k01);
                w = w + W3D(i01 - im + 1, 1, k01 - km + 1);
              }
        } else if (i0 > 1 && i0 <= nx1 - 1) {
          for (i01 = im; i01 <= ip; i01++)
            if (CH(i01, j0, k0) != ex) {
              sum = sum + W3D(i01 - im + 1, 1, 1) * U1(i01, j0, k0);
              w = w + W3D(i01 - im + 1, 1, 1);
            }
        } else if (j0 > 1 && j0 <= ny1 - 1) {
          for (j01 = jm; j01 <= jp; j01++)
            if (CH(i0, j01, k0) != ex) {
              sum = sum + W3D(1, j01 - jm + 1, 1) * U1(i0, j01, k0);
              w = w + W3D(1, j01 - jm + 1, 1);
            }
        } else if (k0 > 1 && k0 <= nz1 - 1) {
          for (k01 = km; k01 <= kp; k01++)
            if (CH(i0, j0, k01) != ex) {
              sum = sum + W3D(1, 1, k01 - km + 1) * U1(i0, j0, k01);
              w = w + w3d_kbranch(i, k01 - km + 1);
            }
        } else {
          sum = U1(i0, j0, k0);
          w = 1;
        }
        if (w == 0 || CH(i0, j0, k0) == ex)
          F3(u2, i2 + (i - 1) * si2, j2 + (j - 1) * sj2, k2 + (k - 1) * sk2, nx2, ny2) = 0;
        else
          F3(u2, i2 + (i - 1) * si2, j2 + (j - 1) * sj2, k2 + (k - 1) * sk2, nx2, ny2) = sum / w;
      }
#undef U1
#undef CH
}

/* src/m_util_r8.f:733-830 */
void or_dmfwr3d(const double *u1, double *u2, int nx1, int ny1, int nz1,
                int nx2, int ny2, int nz2, int i1, int j1, int k1,
                int i2, int j2, int k2, int ni, int nj, int nk,
                int si1, int sj1, int sk1, int si2, int sj2, int sk2,
                const double *chr1, double ex, int do_ex)
{
  int i, j, k, i0, j0, k0, im, jm, km, ip, jp, kp;
  if (i1 < 1 || (i1 + (ni - 1) * si1) > nx1) return;
  if (i2 < 1 || (i2 + (ni - 1) * si2) > nx2) return;
  if (j1 < 1 || (j1 + (nj - 1) * sj1) > ny1) return;
  if (j2 < 1 || (j2 + (nj - 1) * sj2) > ny2) return;
  if (k1 < 1 || (k1 + (nk - 1) * sk1) > nz1) return;
  if (k2 < 1 || (k2 + (nk - 1) * sk2) > nz2) return;
  if (do_ex == 1) {
    dmfwr3d_ex(u1, u2, nx1, ny1, nz1, nx2, ny2, nz2, i1, j1, k1, i2, j2, k2, ni, nj, nk,
               si1, sj1, sk1, si2, sj2, sk2, chr1, ex);
    return;
  }
#define U1(a, b, c) F3(u1, a, b, c, nx1, ny1)
  for (i = 1; i <= ni; i++)
    for (j = 1; j <= nj; j++)
      for (k = 1; k <= nk; k++) {
        double *d = &F3(u2, i2 + (i - 1) * si2, j2 + (j - 1) * sj2, k2 + (k - 1) * sk2, nx2, ny2);
        i0 = i1 + (i - 1) * si1; j0 = j1 + (j - 1) * sj1; k0 = k1 + (k - 1) * sk1;
        im = i0 - 1; jm = j0 - 1; km = k0 - 1; ip = i0 + 1; jp = j0 + 1; kp = k0 + 1;
        if (i0 > 1 && i0 <= nx1 - 1 && j0 > 1 && j0 <= ny1 - 1 && k0 > 1 && k0 <= nz1 - 1) {
          *d = (U1(im, jm, km) + 2 * U1(i0, jm, km) + U1(ip, jm, km) +
                2 * U1(im, j0, km) + 4 * U1(i0, j0, km) + 2 * U1(ip, j0, km) +
                U1(im, jp, km) + 2 * U1(i0, jp, km) + U1(ip, jp, km) +
                2 * U1(im, jm, k0) + 4 * U1(i0, jm, k0) + 2 * U1(ip, jm, k0) +
                4 * U1(im, j0, k0) + 8 * U1(i0, j0, k0) + 4 * U1(ip, j0, k0) +
                2 * U1(im, jp, k0) + 4 * U1(i0, jp, k0) + 2 * U1(ip, jp, k0) +
                U1(im, jm, kp) + 2 * U1(i0, jm, kp) + U1(ip, jm, kp) +
                2 * U1(im, j0, kp) + 4 * U1(i0, j0, kp) + 2 * U1(ip, j0, kp) +
                U1(im, jp, kp) + 2 * U1(i0, jp, kp) + U1(ip, jp, kp)) / 64.0;
        } else if (i0 > 1 && i0 <= nx1 - 1 && j0 > 1 && j0 <= ny1 - 1) {
          *d = (1 * U1(im, jm, k0) + 2 * U1(i0, jm, k0) + 1 * U1(ip, jm, k0) +
                2 * U1(im, j0, k0) + 4 * U1(i0, j0, k0) + 2 * U1(ip, j0, k0) +
                1 * U1(im, jp, k0) + 2 * U1(i0, jp, k0) + 1 * U1(ip, jp, k0)) / 16.0;
        } else if (j0 > 1 && j0 <= ny1 - 1 && k0 > 1 && k0 <= nz1 - 1) {
          *d = (1 * U1(i0, jm, km) + 2 * U1(i0, jm, k0) + 1 * U1(i0, jm, kp) +
                2 * U1(i0, j0, km) + 4 * U1(i0, j0, k0) + 2 * U1(i0, j0, kp) +
                1 * U1(i0, jp, km) + 2 * U1(i0, jp, k0) + 1 * U1(i0, jp, kp)) / 16.0;
        } else if (i0 > 1 && i0 <= nx1 - 1 && k0 > 1 && k0 <= nz1 - 1) {
          *d = (1 * U1(im, j0, km) + 2 * U1(im, j0, k0) + 1 * U1(im, j0, kp) +
                2 * U1(i0, j0, km) + 4 * U1(i0, j0, k0) + 2 * U1(i0, j0, kp) +
                1 * U1(ip, j0, km) + 2 * U1(ip, j0, k0) + 1 * U1(ip, j0, kp)) / 16.0;
        } else if (i0 > 1 && i0 <= nx1 - 1) {
          *d = (U1(im, j0, k0) + 2 * U1(i0, j0, k0) + U1(ip, j0, k0)) / 4.0;
        } else if (j0 > 1 && j0 <= ny1 - 1) {
          *d = (U1(i0, jm, k0) + 2 * U1(i0, j0, k0) + U1(i0, jp, k0)) / 4.0;
        } else if (k0 > 1 && k0 <= nz1 - 1) {
          *d = (U1(i0, j0, km) + 2 * U1(i0, j0, k0) + U1(i0, j0, kp)) / 4.0;
        } else {
          *d = U1(i0, j0, k0);
        }
      }
#undef U1
}

/* ======================================================================== */
/* hooks                                                                     */
/* ======================================================================== */

#define IX(i, j, k) ((size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)ny * (size_t)(k)))

/* Pointwise operator at 0-based interior (i,j,k).  op 0 follows
 * w/examples/nps/num_w.f:42-45 term by term; op 1 is the project's
 * variable-coefficient operator with arithmetic-mean face coefficients. */
static inline double op_apply(int op, const double *V, const double *rho,
                              const double *a, const double *c, int nx, int ny,
                              size_t p, double dx2, double dy2, double dz2,
                              double *jac)
{
  const size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;
  if (op == 0) {
    double L = (V[p + sx] - 2 * V[p] + V[p - sx]) / dx2 +
               (V[p + sy] - 2 * V[p] + V[p - sy]) / dy2 +
               (V[p + sz] - 2 * V[p] + V[p - sz]) / dz2 - rho[p];
    if (jac) *jac = -2 / dx2 - 2 / dy2 - 2 / dz2;
    return L;
  } else {
    double ae = 0.5 * (a[p] + a[p + sx]), aw = 0.5 * (a[p] + a[p - sx]);
    double an = 0.5 * (a[p] + a[p + sy]), as = 0.5 * (a[p] + a[p - sy]);
    double at = 0.5 * (a[p] + a[p + sz]), ab = 0.5 * (a[p] + a[p - sz]);
    double lx = (ae * (V[p + sx] - V[p]) - aw * (V[p] - V[p - sx])) / dx2;
    double ly = (an * (V[p + sy] - V[p]) - as * (V[p] - V[p - sy])) / dy2;
    double lz = (at * (V[p + sz] - V[p]) - ab * (V[p] - V[p - sz])) / dz2;
    double L = lx + ly + lz + c[p] * V[p] - rho[p];
    if (jac) *jac = -(ae + aw) / dx2 - (an + as) / dy2 - (at + ab) / dz2 + c[p];
    return L;
  }
}

static inline int point_on(const double *cmask, const double *chr, double ex, size_t p)
{
  if (cmask[p] != 1.0) return 0;
  if (chr && chr[p] == ex) return 0;
  return 1;
}

/* w/examples/nps/num_w.f:18-52 (loop order k, j, i) */
void or_lop(int op, double *LV, const double *V, const double *rho,
            const double *a, const double *c, const double *cmask,
            const double *chr, double ex, int nx, int ny, int nz,
            double dx, double dy, double dz)
{
  int i, j, k;
  double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
  for (k = 1; k < nz - 1; k++)
    for (j = 1; j < ny - 1; j++)
      for (i = 1; i < nx - 1; i++) {
        size_t p = IX(i, j, k);
        if (point_on(cmask, chr, ex, p)) LV[p] = op_apply(op, V, rho, a, c, nx, ny, p, dx2, dy2, dz2, 0);
      }
}

/* w/examples/nps/num_w.f:59-90 */
double or_residual(int op, double *res, const double *rhs, const double *V,
                   const double *rho, const double *a, const double *c,
                   const double *cmask, const double *chr, double ex,
                   int nx, int ny, int nz, double dx, double dy, double dz)
{
  int i, j, k, sum = 0;
  double norm = 0;
  or_lop(op, res, V, rho, a, c, cmask, chr, ex, nx, ny, nz, dx, dy, dz);
  for (k = 1; k < nz - 1; k++)
    for (j = 1; j < ny - 1; j++)
      for (i = 1; i < nx - 1; i++) {
        size_t p = IX(i, j, k);
        if (point_on(cmask, chr, ex, p)) {
          res[p] = res[p] - rhs[p];
          norm = norm + res[p] * res[p];
          sum = sum + 1;
        }
      }
  return sqrt(norm / sum);
}

/* w/examples/nps/num_w.f:101-157 (loop order pass, i, j, k) */
double or_relax(int op, double *V, const double *rhs, const double *rho,
                const double *a, const double *c, const double *cmask,
                const double *chr, double ex, int nx, int ny, int nz,
                double dx, double dy, double dz)
{
  int pass, i, j, k, sum = 0;
  double norm = 0, dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
  for (pass = 0; pass <= 1; pass++)
    for (i = 1; i < nx - 1; i++)
      for (j = 1; j < ny - 1; j++)
        for (k = 1; k < nz - 1; k++) {
          size_t p = IX(i, j, k);
          /* Fortran indices are 1-based: mod(i+j+k+pass,2) with i,j,k+1 each */
          if (((i + j + k + 3 + pass) % 2) == 0 && point_on(cmask, chr, ex, p)) {
            double jac;
            double res = op_apply(op, V, rho, a, c, nx, ny, p, dx2, dy2, dz2, &jac) - rhs[p];
            V[p] = V[p] - res / jac;
            norm = norm + res * res;
            sum = sum + 1;
          }
        }
  return sqrt(norm / sum);
}

/* ======================================================================== */
/* masks and bookkeeping                                                     */
/* ======================================================================== */

/* amrd/mg.c:432-488 (one grid) */
void or_set_cmask_bdy(double *mask, const int *shape, const double *gbbox,
                      const double *base_bbox, const double *gdx,
                      const int *periodic)
{
  int is, js, ks, ie, je, ke, i, j, k;
  size_t n = (size_t)shape[0] * shape[1] * shape[2], q;
  for (q = 0; q < n; q++) mask[q] = 0.0;
  is = js = ks = 0;
  ie = shape[0] - 1;
  if ((gbbox[0] - base_bbox[0]) > gdx[0] / 2 || periodic[0]) is++;
  if ((base_bbox[1] - gbbox[1]) > gdx[0] / 2 || periodic[0]) ie--;
  je = shape[1] - 1;
  if ((gbbox[2] - base_bbox[2]) > gdx[1] / 2 || periodic[1]) js++;
  if ((base_bbox[3] - gbbox[3]) > gdx[1] / 2 || periodic[1]) je--;
  ke = shape[2] - 1;
  if ((gbbox[4] - base_bbox[4]) > gdx[2] / 2 || periodic[2]) ks++;
  if ((base_bbox[5] - gbbox[5]) > gdx[2] / 2 || periodic[2]) ke--;
  if (is >= 0 && js >= 0 && ks >= 0 && is <= ie && js <= je && ks <= ke &&
      ie < shape[0] && je < shape[1] && ke < shape[2])
    for (i = is; i <= ie; i++)
      for (j = js; j <= je; j++)
        for (k = ks; k <= ke; k++) mask[i + j * (size_t)shape[0] + k * (size_t)shape[0] * shape[1]] = 1.0;
}

/* amrd/mg.c:362-426 (one grid, a list of child bboxes).  The C code assigns
 * doubles to ints, i.e. truncation toward zero, as reproduced here. */
void or_set_cmask_child(double *mask, const int *shape, const double *gbbox,
                        const double *base_bbox, const double *gdx,
                        const int *periodic, int nchild,
                        const double *child_bboxes)
{
  int is, js, ks, ie, je, ke, i, j, k, m;
  size_t n = (size_t)shape[0] * shape[1] * shape[2], q;
  for (q = 0; q < n; q++) mask[q] = 0.0;
  for (m = 0; m < nchild; m++) {
    const double *cb = child_bboxes + 6 * m;
    is = (int)MAX((cb[0] - gbbox[0]) / gdx[0] + 0.5, 0);
    ie = (int)MIN((cb[1] - gbbox[0]) / gdx[0] + 0.5, shape[0] - 1);
    if (is > 0 || periodic[0] || (gbbox[0] - base_bbox[0]) > gdx[0] / 2) is++;
    if (ie < (shape[0] - 1) || periodic[0] || (base_bbox[1] - gbbox[1]) > gdx[0] / 2) ie--;
    js = (int)MAX((cb[2] - gbbox[2]) / gdx[1] + 0.5, 0);
    je = (int)MIN((cb[3] - gbbox[2]) / gdx[1] + 0.5, shape[1] - 1);
    if (js > 0 || periodic[1] || (gbbox[2] - base_bbox[2]) > gdx[1] / 2) js++;
    if (je < (shape[1] - 1) || periodic[1] || (base_bbox[3] - gbbox[3]) > gdx[1] / 2) je--;
    ks = (int)MAX((cb[4] - gbbox[4]) / gdx[2] + 0.5, 0);
    ke = (int)MIN((cb[5] - gbbox[4]) / gdx[2] + 0.5, shape[2] - 1);
    if (ks > 0 || periodic[2] || (gbbox[4] - base_bbox[4]) > gdx[2] / 2) ks++;
    if (ke < (shape[2] - 1) || periodic[2] || (base_bbox[5] - gbbox[5]) > gdx[2] / 2) ke--;
    if (is >= 0 && js >= 0 && ks >= 0 && is <= ie && js <= je && ks <= ke &&
        ie < shape[0] && je < shape[1] && ke < shape[2])
      for (i = is; i <= ie; i++)
        for (j = js; j <= je; j++)
          for (k = ks; k <= ke; k++) mask[i + j * (size_t)shape[0] + k * (size_t)shape[0] * shape[1]] = 1.0;
  }
}

/* src/misc.c:29-32 */
static int fuzz_gt(double x1, double x2, double fuzz) { return x1 > fuzz + x2; }

/* src/gh.c:1427-1467, including its overlap test as written */
void or_calc_island_no(const double *bboxes, int num, double *island_no,
                       const double *dx, int cf, int dim, const int *min_cwidth)
{
  int i, j, k, delta = 1, shape[3], overlap;
  for (i = 0; i < num; i++) {
    island_no[i] = i + 1;
    for (j = 0; j < dim; j++) {
      shape[j] = (int)((bboxes[2 * i * dim + 2 * j + 1] - bboxes[2 * i * dim + 2 * j]) / dx[j] + 1.5);
      if ((((shape[j] - 1) % cf) && ((shape[j] - 1) / cf + 1) >= min_cwidth[j]) ||
          (((shape[j] - 1) / cf + 1) < min_cwidth[j]))
        island_no[i] = 0;
      else if (((shape[j] - 1) / cf / 2 + 1) < min_cwidth[j])
        island_no[i] = -fabs(island_no[i]);
    }
  }
  while (delta) {
    delta = 0;
    for (i = 0; i < num; i++)
      for (j = i + 1; j < num; j++)
        if (island_no[i] != island_no[j]) {
          overlap = 1;
          for (k = 0; k < dim; k++)
            if (fuzz_gt(bboxes[2 * i * dim + 2 * k], bboxes[2 * j * dim + 2 * k + 1], dx[k] / 2) ||
                fuzz_gt(bboxes[2 * i * dim + 2 * k + 1], bboxes[2 * j * dim + 2 * k], dx[k] / 2))
              overlap = 0;
          if (overlap) {
            delta++;
            if (island_no[i] == 0 || island_no[j] == 0)
              island_no[i] = island_no[j] = 0;
            else
              island_no[i] = island_no[j] = MIN(island_no[i], island_no[j]);
          }
        }
  }
}

/* src/transfer.c:73-95 */
void or_calc_gs_ibbox(const double *seg_bbox, const double *g_bbox,
                      const int *g_shape, int dim, int *ibbox, double *dx)
{
  int i;
  for (i = 0; i < dim; i++) {
    if (g_shape[i] > 1) {
      dx[i] = (g_bbox[2 * i + 1] - g_bbox[2 * i]) / (g_shape[i] - 1);
      ibbox[2 * i] = (int)((seg_bbox[2 * i] - g_bbox[2 * i]) / dx[i] + 1.01);
      ibbox[2 * i + 1] = (int)((seg_bbox[2 * i + 1] - g_bbox[2 * i]) / dx[i] + 1.99);
    } else {
      ibbox[2 * i] = 1;
      ibbox[2 * i + 1] = 1;
      dx[i] = 1e-20;
    }
  }
}

/* ======================================================================== */
/* uniform-hierarchy FAS driver                                             */
/* ======================================================================== */

enum { F_U = 0, F_RHO, F_RHS, F_RES, F_LOP, F_CMASK, F_A, F_C, F_CHR, F_NUM };

struct or_mg {
  int nlev, op, do_ex;
  double ex;
  int shape[OR_MAXLEV][3];
  double bbox[6], base_bbox[6];
  double hdx[OR_MAXLEV][3]; /* hook dx: x(2)-x(1) */
  double gdx[OR_MAXLEV][3]; /* mg_ldptr dx: x[2]-x[1] (amrd/mg.c:220) */
  double *f[OR_MAXLEV][F_NUM];
  int pre, pst, max_citer, rop_u, rop_res, rop_src, iord;
  double crtol, w0_r, w0_i;
  int periodic[3];
};

static size_t lev_size(const or_mg *m, int L)
{
  return (size_t)m->shape[L - 1][0] * m->shape[L - 1][1] * m->shape[L - 1][2];
}

/* Coordinates by repeated addition (src/gh.c:905-906) and the two dx
 * conventions used downstream. */
static void level_dx(or_mg *m, int L)
{
  int d;
  for (d = 0; d < 3; d++) {
    int n = m->shape[L - 1][d];
    double dx = (m->bbox[2 * d + 1] - m->bbox[2 * d]) / (n - 1);
    double x0 = m->bbox[2 * d], x1 = x0 + dx, x2 = x1 + dx;
    m->hdx[L - 1][d] = x1 - x0;
    m->gdx[L - 1][d] = x2 - x1;
  }
}

or_mg *or_mg_create(int nx, int ny, int nz, const double *bbox, int mcw,
                    int op, int do_ex, double ex)
{
  or_mg *m = (or_mg *)calloc(1, sizeof(or_mg));
  int shapes[OR_MAXLEV][3], nl = 0, cf, d, L, w;
  double dx0[3], island;
  int mcwv[3] = {mcw, mcw, mcw};
  memcpy(m->bbox, bbox, 6 * sizeof(double));
  memcpy(m->base_bbox, bbox, 6 * sizeof(double));
  m->op = op;
  m->do_ex = do_ex;
  m->ex = ex;
  shapes[0][0] = nx; shapes[0][1] = ny; shapes[0][2] = nz;
  nl = 1;
  for (d = 0; d < 3; d++) dx0[d] = (bbox[2 * d + 1] - bbox[2 * d]) / (shapes[0][d] - 1);
  /* build_mgh (src/gh.c:1275-1340): coarsen by 2 while the island number
   * of the base box is nonzero */
  for (cf = 2;; cf *= 2) {
    or_calc_island_no(bbox, 1, &island, dx0, cf, 3, mcwv);
    if (island == 0 || nl >= OR_MAXLEV) break;
    for (d = 0; d < 3; d++) {
      double ldx = dx0[d] * cf;
      shapes[nl][d] = (int)((bbox[2 * d + 1] - bbox[2 * d]) / ldx + 1.5);
    }
    nl++;
  }
  m->nlev = nl;
  for (L = 1; L <= nl; L++) {
    for (d = 0; d < 3; d++) m->shape[L - 1][d] = shapes[nl - L][d];
    level_dx(m, L);
    for (w = 0; w < F_NUM; w++) {
      if ((w == F_A || w == F_C) && op == 0) continue;
      if (w == F_CHR && !do_ex) continue;
      m->f[L - 1][w] = (double *)calloc(lev_size(m, L), sizeof(double));
    }
  }
  m->pre = m->pst = 2;
  m->max_citer = 5000;
  m->crtol = 1e-3;
  m->w0_r = m->w0_i = 1.0;
  m->rop_u = m->rop_res = m->rop_src = 3;
  m->iord = 2;
  return m;
}

void or_mg_destroy(or_mg *m)
{
  int L, w;
  if (!m) return;
  for (L = 0; L < m->nlev; L++)
    for (w = 0; w < F_NUM; w++) free(m->f[L][w]);
  free(m);
}

int or_mg_nlev(const or_mg *m) { return m->nlev; }
void or_mg_shape(const or_mg *m, int lev, int *shape)
{
  shape[0] = m->shape[lev - 1][0];
  shape[1] = m->shape[lev - 1][1];
  shape[2] = m->shape[lev - 1][2];
}
double *or_mg_field(or_mg *m, int lev, int which)
{
  if (lev < 1 || lev > m->nlev || which < 0 || which >= F_NUM) return 0;
  return m->f[lev - 1][which];
}

void or_mg_set_params(or_mg *m, int pre, int pst, int max_citer,
                      double crtol, double w0_r, double w0_i, int rop_u,
                      int rop_res, int rop_src, int interp_ord)
{
  m->pre = pre; m->pst = pst; m->max_citer = max_citer; m->crtol = crtol;
  m->w0_r = w0_r; m->w0_i = w0_i; m->rop_u = rop_u; m->rop_res = rop_res;
  m->rop_src = rop_src; m->iord = interp_ord;
}

/* local_transfer coarsen branch (src/transfer.c:1126-1149) for one field,
 * whole fine grid Lf onto whole coarse grid Lf-1 (single grid per level:
 * calc_gs_ibbox gives [1,n] on both, nib = coarse count, stride 2). */
static void restrict_field(or_mg *m, int Lfine, int w, int op)
{
  const int *sf = m->shape[Lfine - 1], *sc = m->shape[Lfine - 2];
  const double *src = m->f[Lfine - 1][w];
  double *dst = m->f[Lfine - 2][w];
  /* the coefficients of the variable-coefficient operator (a project-defined
   * hook with no reference counterpart) have no excised value: they are
   * restricted without the excision renormalisation */
  const int do_ex = m->do_ex && w != F_A && w != F_C;
  const double *chr = do_ex ? m->f[Lfine - 1][F_CHR] : src;
  if (!src || !dst || op == 0) return;
  if (op == 1)
    or_dmcopy3d(src, dst, sf[0], sf[1], sf[2], sc[0], sc[1], sc[2], 1, 1, 1, 1, 1, 1, sc[0], sc[1], sc[2],
                2, 2, 2, 1, 1, 1);
  else if (op == 2)
    or_dmhwr3d(src, dst, sf[0], sf[1], sf[2], sc[0], sc[1], sc[2], 1, 1, 1, 1, 1, 1, sc[0], sc[1], sc[2],
               2, 2, 2, 1, 1, 1, chr, m->ex, do_ex);
  else
    or_dmfwr3d(src, dst, sf[0], sf[1], sf[2], sc[0], sc[1], sc[2], 1, 1, 1, 1, 1, 1, sc[0], sc[1], sc[2],
               2, 2, 2, 1, 1, 1, chr, m->ex, do_ex);
}

/* PAMR_inject(L+1 -> L) on a pure MG level: TF_INJECT_TO_MG_LEV selects
 * every gfn with mg_inject (src/pamr.c:677-681,1886-1897). */
static void inject(or_mg *m, int Lfine)
{
  restrict_field(m, Lfine, F_U, m->rop_u);
  restrict_field(m, Lfine, F_RES, m->rop_res);
  restrict_field(m, Lfine, F_RHO, m->rop_src);
  if (m->op == 1) {
    restrict_field(m, Lfine, F_A, m->rop_src);
    restrict_field(m, Lfine, F_C, m->rop_src);
  }
}

void or_mg_init(or_mg *m)
{
  int L;
  /* TF_MGH_INIT (src/gh.c:1237-1239,1334): in_amrh gfns with mg_inject */
  for (L = m->nlev; L > 1; L--) {
    restrict_field(m, L, F_U, m->rop_u);
    restrict_field(m, L, F_RHO, m->rop_src);
    if (m->op == 1) {
      restrict_field(m, L, F_A, m->rop_src);
      restrict_field(m, L, F_C, m->rop_src);
    }
  }
}

/* amrd/mg.c:242-262 */
static void clear_mg_vars(or_mg *m, int L)
{
  size_t n = lev_size(m, L), j;
  double *res = m->f[L - 1][F_RES], *rhs = m->f[L - 1][F_RHS], *lop = m->f[L - 1][F_LOP];
  for (j = 0; j < n; j++) res[j] = rhs[j] = lop[j] = 0;
}

/* amrd/mg.c:271-314 (save_tre = 0) */
static void compute_mg_rhs(or_mg *m, int L)
{
  size_t n = lev_size(m, L), j;
  double *res = m->f[L - 1][F_RES], *rhs = m->f[L - 1][F_RHS], *lop = m->f[L - 1][F_LOP];
  const double *u = m->f[L - 1][F_U], *cm = m->f[L - 1][F_CMASK];
  for (j = 0; j < n; j++) {
    if (cm[j] == 1.0)
      rhs[j] = -res[j] * m->w0_r + lop[j];
    else
      rhs[j] = 0;
    lop[j] = u[j];
  }
}

/* PAMR_interp(Lc -> Lc+1) of f_res on a single grid per level: the whole
 * coarse grid is interpolated into a fine-size temporary and copied back
 * (src/transfer.c:1007-1051,1089-1115, interp_buffer = 2 covers all). */
static void interp_res(or_mg *m, int Lc)
{
  const int *sc = m->shape[Lc - 1], *sf = m->shape[Lc];
  double *src = m->f[Lc - 1][F_RES], *dst = m->f[Lc][F_RES];
  size_t nf = lev_size(m, Lc + 1);
  double *tmp = (double *)malloc(nf * sizeof(double));
  const double *chr1 = m->do_ex ? m->f[Lc - 1][F_CHR] : src;
  or_dminterp3d(src, tmp, sc[0], sc[1], sc[2], sf[0], sf[1], sf[2], 1, 1, 1, 1, 1, 1, sc[0], sc[1], sc[2],
                1, 1, 1, 2, 2, 2, m->iord, chr1, m->ex, m->do_ex);
  or_dmcopy3d(tmp, dst, sf[0], sf[1], sf[2], sf[0], sf[1], sf[2], 1, 1, 1, 1, 1, 1, sf[0], sf[1], sf[2],
              1, 1, 1, 1, 1, 1);
  free(tmp);
  if (m->do_ex) {
    /* clear_ex (src/transfer.c:930-935,1110-1114) */
    size_t nc = lev_size(m, Lc), q;
    const double *cc = m->f[Lc - 1][F_CHR], *cfch = m->f[Lc][F_CHR];
    for (q = 0; q < nc; q++) if (cc[q] == m->ex) src[q] = 0;
    for (q = 0; q < nf; q++) if (cfch[q] == m->ex) dst[q] = 0;
  }
}

/* amrd/mg.c:323-355 */
static void compute_cgc(or_mg *m, int L)
{
  size_t n = lev_size(m, L - 1), j;
  double *res = m->f[L - 2][F_RES];
  const double *u = m->f[L - 2][F_U], *lop = m->f[L - 2][F_LOP];
  for (j = 0; j < n; j++) res[j] = (u[j] - lop[j]) / m->w0_i;
  interp_res(m, L - 1);
  n = lev_size(m, L);
  {
    double *uf = m->f[L - 1][F_U];
    const double *rf = m->f[L - 1][F_RES];
    for (j = 0; j < n; j++) uf[j] += rf[j];
  }
}

static const double *chr_of(or_mg *m, int L) { return m->do_ex ? m->f[L - 1][F_CHR] : 0; }

static double hook_relax(or_mg *m, int L)
{
  const int *s = m->shape[L - 1];
  double **f = m->f[L - 1];
  return or_relax(m->op, f[F_U], f[F_RHS], f[F_RHO], f[F_A], f[F_C], f[F_CMASK], chr_of(m, L), m->ex,
                  s[0], s[1], s[2], m->hdx[L - 1][0], m->hdx[L - 1][1], m->hdx[L - 1][2]);
}
static double hook_residual(or_mg *m, int L)
{
  const int *s = m->shape[L - 1];
  double **f = m->f[L - 1];
  return or_residual(m->op, f[F_RES], f[F_RHS], f[F_U], f[F_RHO], f[F_A], f[F_C], f[F_CMASK], chr_of(m, L),
                     m->ex, s[0], s[1], s[2], m->hdx[L - 1][0], m->hdx[L - 1][1], m->hdx[L - 1][2]);
}
static void hook_lop(or_mg *m, int L)
{
  const int *s = m->shape[L - 1];
  double **f = m->f[L - 1];
  or_lop(m->op, f[F_LOP], f[F_U], f[F_RHO], f[F_A], f[F_C], f[F_CMASK], chr_of(m, L), m->ex, s[0], s[1], s[2],
         m->hdx[L - 1][0], m->hdx[L - 1][1], m->hdx[L - 1][2]);
}
static void cmask_bdy(or_mg *m, int L)
{
  or_set_cmask_bdy(m->f[L - 1][F_CMASK], m->shape[L - 1], m->bbox, m->base_bbox, m->gdx[L - 1], m->periodic);
}
static void cmask_child(or_mg *m, int L)
{
  /* the single child grid at L+1 covers the whole domain */
  or_set_cmask_child(m->f[L - 1][F_CMASK], m->shape[L - 1], m->bbox, m->base_bbox, m->gdx[L - 1], m->periodic,
                     1, m->bbox);
}

/* amrd/mg.c:84-201 */
double or_mg_vcycle(or_mg *m, double *lev_norms, int *citer)
{
  int L, Lf = m->nlev, iter;
  double res[OR_MAXLEV + 2], ctol;
  for (L = 0; L < OR_MAXLEV + 2; L++) res[L] = 0;
  clear_mg_vars(m, Lf);
  for (L = Lf; L > 0; L--) {
    if (L < Lf) {
      res[L + 1] = hook_residual(m, L + 1);
      /* PAMR_sync on one grid per level is a no-op (src/transfer.c:527) */
      clear_mg_vars(m, L);
      inject(m, L + 1);
      cmask_child(m, L);
      hook_lop(m, L);
      compute_mg_rhs(m, L);
    }
    cmask_bdy(m, L);
    for (iter = 0; iter < m->pre; iter++) res[L] = hook_relax(m, L);
  }
  for (L = 1; L <= Lf; L++) {
    ctol = res[L] * m->crtol;
    if (L > 1) compute_cgc(m, L);
    for (iter = 0; iter < m->pst; iter++) res[L] = hook_relax(m, L);
    if (L == 1) { /* only level 1 is flagged coarsest (src/gh.c:1353-1354) */
      res[L] = ctol + 1;
      while (iter < m->max_citer && res[L] > ctol) {
        res[L] = hook_relax(m, L);
        iter++;
      }
      if (citer) *citer = iter;
    }
    if (L == Lf) res[Lf] = hook_residual(m, L);
  }
  if (lev_norms)
    for (L = 1; L <= Lf; L++) lev_norms[L - 1] = res[L];
  return res[Lf];
}

/* amrd/mg.c:56-72 */
int or_mg_solve(or_mg *m, int min_iter, int max_iter, double tol, double *norms_out)
{
  int it = 0;
  double res = tol + 1;
  while ((it < min_iter || res > tol) && it < max_iter) {
    res = or_mg_vcycle(m, 0, 0);
    if (norms_out) norms_out[it] = res;
    it++;
  }
  return it;
}

double or_mg_residual_fine(or_mg *m)
{
  int Lf = m->nlev;
  cmask_bdy(m, Lf);
  return hook_residual(m, Lf);
}

/* ======================================================================== */
/* learned smoother network oracle                                           */
/* ======================================================================== */

static double gelu_t(double x)
{
  const double k = 0.7978845608028654; /* sqrt(2/pi) */
  double u = k * (x + 0.044715 * x * x * x);
  return 0.5 * x * (1.0 + tanh(u));
}
static double gelu_t_grad(double x)
{
  const double k = 0.7978845608028654;
  double x2 = x * x;
  double u = k * (x + 0.044715 * x2 * x);
  double t = tanh(u);
  return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * k * (1.0 + 3.0 * 0.044715 * x2);
}

/* y[co] = b[co] + sum_ci sum_tap w[co][ci][tap] x[ci](p + off(tap)), zero padded */
static void conv3(int cin, int cout, const double *w, const double *b, const double *x, double *y,
                  int n0, int n1, int n2)
{
  size_t np = (size_t)n0 * n1 * n2;
  int co, ci, i, j, k, t;
  for (co = 0; co < cout; co++)
    for (k = 0; k < n2; k++)
      for (j = 0; j < n1; j++)
        for (i = 0; i < n0; i++) {
          double s = b[co];
          for (ci = 0; ci < cin; ci++)
            for (t = 0; t < 27; t++) {
              int di = t % 3 - 1, dj = (t / 3) % 3 - 1, dk = t / 9 - 1;
              int ii = i + di, jj = j + dj, kk = k + dk;
              if (ii < 0 || jj < 0 || kk < 0 || ii >= n0 || jj >= n1 || kk >= n2) continue;
              s += w[((size_t)co * cin + ci) * 27 + t] *
                   x[(size_t)ci * np + ii + (size_t)n0 * (jj + (size_t)n1 * kk)];
            }
          y[(size_t)co * np + i + (size_t)n0 * (j + (size_t)n1 * k)] = s;
        }
}

/* gradients of conv3: gx (may be NULL) += W^T gy, gw += gy x^T, gb += sum gy */
static void conv3_back(int cin, int cout, const double *w, const double *x, const double *gy, double *gx,
                       double *gw, double *gb, int n0, int n1, int n2)
{
  size_t np = (size_t)n0 * n1 * n2;
  int co, ci, i, j, k, t;
  for (co = 0; co < cout; co++)
    for (k = 0; k < n2; k++)
      for (j = 0; j < n1; j++)
        for (i = 0; i < n0; i++) {
          double g = gy[(size_t)co * np + i + (size_t)n0 * (j + (size_t)n1 * k)];
          gb[co] += g;
          for (ci = 0; ci < cin; ci++)
            for (t = 0; t < 27; t++) {
              int di = t % 3 - 1, dj = (t / 3) % 3 - 1, dk = t / 9 - 1;
              int ii = i + di, jj = j + dj, kk = k + dk;
              size_t q;
              if (ii < 0 || jj < 0 || kk < 0 || ii >= n0 || jj >= n1 || kk >= n2) continue;
              q = (size_t)ci * np + ii + (size_t)n0 * (jj + (size_t)n1 * kk);
              gw[((size_t)co * cin + ci) * 27 + t] += g * x[q];
              if (gx) gx[q] += g * w[((size_t)co * cin + ci) * 27 + t];
            }
        }
}

void or_net_forward(int H, const double *w1, const double *b1,
                    const double *w2, const double *b2, const double *w3,
                    const double *b3, const double *x, double *y,
                    int n0, int n1, int n2)
{
  size_t np = (size_t)n0 * n1 * n2, q;
  double *h1 = (double *)malloc(H * np * sizeof(double));
  double *h2 = (double *)malloc(H * np * sizeof(double));
  conv3(1, H, w1, b1, x, h1, n0, n1, n2);
  for (q = 0; q < H * np; q++) h1[q] = gelu_t(h1[q]);
  conv3(H, H, w2, b2, h1, h2, n0, n1, n2);
  for (q = 0; q < H * np; q++) h2[q] = gelu_t(h2[q]);
  conv3(H, 1, w3, b3, h2, y, n0, n1, n2);
  free(h1);
  free(h2);
}

double or_net_loss_grad(int H, const double *w1, const double *b1,
                        const double *w2, const double *b2, const double *w3,
                        const double *b3, const double *x, const double *t,
                        int n0, int n1, int n2, double *gw1, double *gb1,
                        double *gw2, double *gb2, double *gw3, double *gb3)
{
  size_t np = (size_t)n0 * n1 * n2, q;
  double *z1 = (double *)malloc(H * np * sizeof(double)), *h1 = (double *)malloc(H * np * sizeof(double));
  double *z2 = (double *)malloc(H * np * sizeof(double)), *h2 = (double *)malloc(H * np * sizeof(double));
  double *y = (double *)malloc(np * sizeof(double)), *gy = (double *)malloc(np * sizeof(double));
  double *g2 = (double *)calloc(H * np, sizeof(double)), *g1 = (double *)calloc(H * np, sizeof(double));
  double loss = 0;
  memset(gw1, 0, sizeof(double) * H * 27); memset(gb1, 0, sizeof(double) * H);
  memset(gw2, 0, sizeof(double) * H * H * 27); memset(gb2, 0, sizeof(double) * H);
  memset(gw3, 0, sizeof(double) * H * 27); memset(gb3, 0, sizeof(double));
  conv3(1, H, w1, b1, x, z1, n0, n1, n2);
  for (q = 0; q < H * np; q++) h1[q] = gelu_t(z1[q]);
  conv3(H, H, w2, b2, h1, z2, n0, n1, n2);
  for (q = 0; q < H * np; q++) h2[q] = gelu_t(z2[q]);
  conv3(H, 1, w3, b3, h2, y, n0, n1, n2);
  for (q = 0; q < np; q++) {
    double d = y[q] - t[q];
    loss += d * d;
    gy[q] = d / (double)np;
  }
  loss = 0.5 * loss / (double)np;
  conv3_back(H, 1, w3, h2, gy, g2, gw3, gb3, n0, n1, n2);
  for (q = 0; q < H * np; q++) g2[q] *= gelu_t_grad(z2[q]);
  conv3_back(H, H, w2, h1, g2, g1, gw2, gb2, n0, n1, n2);
  for (q = 0; q < H * np; q++) g1[q] *= gelu_t_grad(z1[q]);
  conv3_back(1, H, w1, x, g1, 0, gw1, gb1, n0, n1, n2);
  free(z1); free(h1); free(z2); free(h2); free(y); free(gy); free(g2); free(g1);
  return loss;
}
